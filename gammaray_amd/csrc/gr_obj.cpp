// gr_obj.cpp -- Wavefront OBJ meshes: the state Mesh::from_obj builds
// (src/geom/mesh.rs:61-188), parsed on the host.
//
//   * `v`, `vt`, `vn`, `f` statements; `o`, `g`, `s`, `usemtl`, `mtllib`, `l`,
//     `p` and comments are read past (lines and points are not triangles, and
//     the reference keeps triangles only, mesh.rs:126).  Indices are the
//     file's 1-based global ones, which is what the reference's per-object
//     offsets (mesh.rs:99-120) add back up to.
//   * Numbers are read as f64 and rounded to f32, as the reference casts the
//     parser's f64 values (mesh.rs:102,110,118).
//   * Vertices go through Xform::transform (src/core/xform.rs:41-47), normals
//     through Xform::transform_normal with the inverse from Mat::inverted
//     (src/core/matrix.rs:154-290) and are normalised, except nearly-zero
//     ones, which are kept as read (mesh.rs:107-114).
//   * A face of n > 3 corners becomes the fan (0, k, k+1), k = 1..n-2.
//   * Per triangle: the normal indices when all three corners carry non-zero
//     normals, else GR_NONE (geometric normal, mesh.rs:138-153); the uv
//     indices when all three carry one, else GR_NONE (the zero uv,
//     mesh.rs:129-137).
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gr_internal.h"

struct gr_obj {
    std::vector<float> verts, normals, uvs;  // 3, 3, 2 floats each
    std::vector<uint32_t> tris;              // 9 per triangle
};

namespace {

#pragma clang fp contract(off)

using Mat4 = float[4][4];

// Mat::inverted (src/core/matrix.rs:154-290): 2x2 determinants of column
// pairs, 3x3 cofactors, one reciprocal of the determinant; a determinant that
// is not positive gives the scale matrix of f32::MAX.
void invert(const Mat4 m, Mat4 out) {
    float y01 = m[0][0] * m[1][1] - m[1][0] * m[0][1];
    float y02 = m[0][0] * m[2][1] - m[2][0] * m[0][1];
    float y03 = m[0][0] * m[3][1] - m[3][0] * m[0][1];
    float y12 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
    float y13 = m[1][0] * m[3][1] - m[3][0] * m[1][1];
    float y23 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
    const float z33 = m[0][2] * y12 - m[1][2] * y02 + m[2][2] * y01;
    const float z23 = m[1][2] * y03 - m[3][2] * y01 - m[0][2] * y13;
    const float z13 = m[0][2] * y23 - m[2][2] * y03 + m[3][2] * y02;
    const float z03 = m[2][2] * y13 - m[3][2] * y12 - m[1][2] * y23;
    const float z32 = m[1][3] * y02 - m[2][3] * y01 - m[0][3] * y12;
    const float z22 = m[0][3] * y13 - m[1][3] * y03 + m[3][3] * y01;
    const float z12 = m[2][3] * y03 - m[3][3] * y02 - m[0][3] * y23;
    const float z02 = m[1][3] * y23 - m[2][3] * y13 + m[3][3] * y12;
    y01 = m[0][2] * m[1][3] - m[1][2] * m[0][3];
    y02 = m[0][2] * m[2][3] - m[2][2] * m[0][3];
    y03 = m[0][2] * m[3][3] - m[3][2] * m[0][3];
    y12 = m[1][2] * m[2][3] - m[2][2] * m[1][3];
    y13 = m[1][2] * m[3][3] - m[3][2] * m[1][3];
    y23 = m[2][2] * m[3][3] - m[3][2] * m[2][3];
    const float z30 = m[1][1] * y02 - m[2][1] * y01 - m[0][1] * y12;
    const float z20 = m[0][1] * y13 - m[1][1] * y03 + m[3][1] * y01;
    const float z10 = m[2][1] * y03 - m[3][1] * y02 - m[0][1] * y23;
    const float z00 = m[1][1] * y23 - m[2][1] * y13 + m[3][1] * y12;
    const float z31 = m[0][0] * y12 - m[1][0] * y02 + m[2][0] * y01;
    const float z21 = m[1][0] * y03 - m[3][0] * y01 - m[0][0] * y13;
    const float z11 = m[0][0] * y23 - m[2][0] * y03 + m[3][0] * y02;
    const float z01 = m[2][0] * y13 - m[3][0] * y12 - m[1][0] * y23;
    const float det = m[3][0] * z30 + m[2][0] * z20 + m[1][0] * z10 + m[0][0] * z00;
    std::memset(out, 0, sizeof(Mat4));
    if (!gr::positive(det)) {  // math::is_positive
        out[0][0] = out[1][1] = out[2][2] = FLT_MAX;
        out[3][3] = 1.0f;
        return;
    }
    const float r = 1.0f / det;
    const float cof[4][4] = {{z00, z10, z20, z30}, {z01, z11, z21, z31}, {z02, z12, z22, z32}, {z03, z13, z23, z33}};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[i][j] = cof[i][j] * r;
}

struct Corner {
    long v, t, n;  // 1-based; 0 = absent
};

// One corner "v", "v/t", "v//n" or "v/t/n".
bool parse_corner(const char*& p, Corner* c) {
    c->v = c->t = c->n = 0;
    char* end = nullptr;
    c->v = std::strtol(p, &end, 10);
    if (end == p) return false;
    p = end;
    if (*p == '/') {
        ++p;
        if (*p != '/') {
            c->t = std::strtol(p, &end, 10);
            if (end == p) return false;
            p = end;
        }
        if (*p == '/') {
            ++p;
            c->n = std::strtol(p, &end, 10);
            if (end == p) return false;
            p = end;
        }
    }
    return *p == '\0' || *p == ' ' || *p == '\t' || *p == '\r' || *p == '\n';
}

int parse_floats(const char* p, double* out, int max) {
    int k = 0;
    while (k < max) {
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '\0' || *p == '\r' || *p == '\n' || *p == '#') break;
        char* end = nullptr;
        out[k] = std::strtod(p, &end);
        if (end == p) return -1;
        p = end;
        ++k;
    }
    return k;
}

bool starts(const char* p, const char* kw) {
    const size_t n = std::strlen(kw);
    return std::strncmp(p, kw, n) == 0 && (p[n] == ' ' || p[n] == '\t');
}

}  // namespace

extern "C" {

int gr_obj_load(const char* path, const float* xf, gr_obj** out, char* err, size_t err_len) {
    auto fail = [&](const std::string& msg) {
        if (err && err_len) std::snprintf(err, err_len, "%s", msg.c_str());
        return GR_ERR_INVALID;
    };
    if (!path || !out) return fail("null argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(std::string("Couldn't open OBJ file: ") + std::strerror(errno));
    Mat4 m, inv;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m[i][j] = xf ? xf[4 * i + j] : (i == j ? 1.0f : 0.0f);
    invert(m, inv);
    gr_obj* o = new (std::nothrow) gr_obj();
    if (!o) {
        std::fclose(f);
        return fail("out of host memory");
    }
    // a face may only name records read before it
    std::string line;
    std::vector<Corner> face;
    std::vector<char> buf(1 << 16);
    long lineno = 0;
    int code = GR_OK;
    std::string msg;
    auto bad = [&](const char* what) {
        code = GR_ERR_INVALID;
        msg = "OBJ parse error at line " + std::to_string(lineno) + ": " + what;
    };
    while (code == GR_OK && std::fgets(buf.data(), (int)buf.size(), f)) {
        line.assign(buf.data());
        while (!line.empty() && line.back() != '\n' && std::fgets(buf.data(), (int)buf.size(), f)) line += buf.data();
        ++lineno;
        const char* p = line.c_str();
        while (*p == ' ' || *p == '\t') ++p;
        double d[4];
        if (starts(p, "v")) {
            const int k = parse_floats(p + 1, d, 4);
            if (k < 3) {
                bad("a vertex needs three coordinates");
                break;
            }
            const float x = (float)d[0], y = (float)d[1], z = (float)d[2];
            // Xform::transform: row vector times the matrix, then the divide by w
            const float tx = x * m[0][0] + y * m[1][0] + z * m[2][0] + m[3][0];
            const float ty = x * m[0][1] + y * m[1][1] + z * m[2][1] + m[3][1];
            const float tz = x * m[0][2] + y * m[1][2] + z * m[2][2] + m[3][2];
            const float tw = x * m[0][3] + y * m[1][3] + z * m[2][3] + m[3][3];
            o->verts.insert(o->verts.end(), {tx / tw, ty / tw, tz / tw});
        } else if (starts(p, "vt")) {
            const int k = parse_floats(p + 2, d, 3);
            if (k < 1) {
                bad("a texture vertex needs a coordinate");
                break;
            }
            o->uvs.insert(o->uvs.end(), {(float)d[0], k > 1 ? (float)d[1] : 0.0f});
        } else if (starts(p, "vn")) {
            const int k = parse_floats(p + 2, d, 3);
            if (k < 3) {
                bad("a normal needs three coordinates");
                break;
            }
            const float x = (float)d[0], y = (float)d[1], z = (float)d[2];
            if (gr::nearly_zero(x * x + y * y + z * z)) {
                o->normals.insert(o->normals.end(), {x, y, z});
            } else {
                // Xform::transform_normal: the transposed inverse, then normalised
                const float nx = x * inv[0][0] + y * inv[0][1] + z * inv[0][2];
                const float ny = x * inv[1][0] + y * inv[1][1] + z * inv[1][2];
                const float nz = x * inv[2][0] + y * inv[2][1] + z * inv[2][2];
                const float len = std::sqrt(nx * nx + ny * ny + nz * nz);
                o->normals.insert(o->normals.end(), {nx / len, ny / len, nz / len});
            }
        } else if (starts(p, "f")) {
            face.clear();
            const char* q = p + 1;
            for (;;) {
                while (*q == ' ' || *q == '\t') ++q;
                if (*q == '\0' || *q == '\r' || *q == '\n' || *q == '#') break;
                Corner c;
                if (!parse_corner(q, &c)) {
                    bad("malformed face corner");
                    break;
                }
                face.push_back(c);
            }
            if (code != GR_OK) break;
            if (face.size() < 3) {
                bad("a face needs three corners");
                break;
            }
            const long nv = (long)(o->verts.size() / 3), nt = (long)(o->uvs.size() / 2),
                       nn = (long)(o->normals.size() / 3);
            for (const Corner& c : face) {
                if (c.v < 1 || c.v > nv || c.t < 0 || c.t > nt || c.n < 0 || c.n > nn) {
                    bad("index out of range (indices are 1-based and name earlier records)");
                    break;
                }
            }
            if (code != GR_OK) break;
            for (size_t k = 1; k + 1 < face.size(); ++k) {
                const Corner* cs[3] = {&face[0], &face[k], &face[k + 1]};
                uint32_t t[9];
                bool has_t = true, has_n = true;
                for (int j = 0; j < 3; ++j) {
                    t[j] = (uint32_t)(cs[j]->v - 1);
                    has_t = has_t && cs[j]->t > 0;
                    has_n = has_n && cs[j]->n > 0;
                }
                if (has_n) {
                    for (int j = 0; j < 3; ++j) {
                        const float* n = &o->normals[3 * (size_t)(cs[j]->n - 1)];
                        if (gr::nearly_zero(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])) has_n = false;
                    }
                }
                for (int j = 0; j < 3; ++j) {
                    t[3 + j] = has_n ? (uint32_t)(cs[j]->n - 1) : GR_NONE;
                    t[6 + j] = has_t ? (uint32_t)(cs[j]->t - 1) : GR_NONE;
                }
                o->tris.insert(o->tris.end(), t, t + 9);
            }
        }
        // anything else (o, g, s, usemtl, mtllib, l, p, comments, blank lines) is read past
    }
    std::fclose(f);
    if (code == GR_OK && o->tris.empty()) {
        code = GR_ERR_INVALID;
        msg = "OBJ file has no triangles";
    }
    if (code == GR_OK && (o->verts.size() / 3 > 0xfffffffeull || o->tris.size() / 9 > 0xffffffffull)) {
        code = GR_ERR_INVALID;
        msg = "OBJ file too large";
    }
    if (code != GR_OK) {
        delete o;
        return fail(msg);
    }
    *out = o;
    return GR_OK;
}

int gr_obj_sizes(const gr_obj* o, uint32_t* nverts, uint32_t* nnormals, uint32_t* nuvs, uint32_t* ntris) {
    if (!o) return GR_ERR_INVALID;
    if (nverts) *nverts = (uint32_t)(o->verts.size() / 3);
    if (nnormals) *nnormals = (uint32_t)(o->normals.size() / 3);
    if (nuvs) *nuvs = (uint32_t)(o->uvs.size() / 2);
    if (ntris) *ntris = (uint32_t)(o->tris.size() / 9);
    return GR_OK;
}

int gr_obj_read(const gr_obj* o, float* verts, float* normals, float* uvs, uint32_t* tris) {
    if (!o) return GR_ERR_INVALID;
    if (verts) std::memcpy(verts, o->verts.data(), o->verts.size() * sizeof(float));
    if (normals) std::memcpy(normals, o->normals.data(), o->normals.size() * sizeof(float));
    if (uvs) std::memcpy(uvs, o->uvs.data(), o->uvs.size() * sizeof(float));
    if (tris) std::memcpy(tris, o->tris.data(), o->tris.size() * sizeof(uint32_t));
    return GR_OK;
}

void gr_obj_destroy(gr_obj* o) { delete o; }

int gr_scene_add_obj(gr_scene* s, const gr_obj* o, const float* material, uint32_t* prim_out) {
    if (!s || !o) return GR_ERR_INVALID;
    return gr_scene_add_mesh(s, o->verts.data(), (uint32_t)(o->verts.size() / 3),
                             o->normals.empty() ? nullptr : o->normals.data(), (uint32_t)(o->normals.size() / 3),
                             o->uvs.empty() ? nullptr : o->uvs.data(), (uint32_t)(o->uvs.size() / 2),
                             o->tris.data(), (uint32_t)(o->tris.size() / 9), material, prim_out);
}

}  // extern "C"
