// gr_render.hip -- wavefront path tracing on gfx950: camera rays, shading and
// bounce loop, next-event estimation, film accumulation.
//
// One render pass is the reference's Stage::trace (src/render/stage.rs:23-36)
// for `spp` iterations: per iteration every pixel gets one jittered sample
// (src/render/film.rs:55-76), the integrator runs per sample, and samples
// splat into the film with the Mitchell filter (film.rs:78-114).  Here the
// samples of several iterations are traced together as one wavefront:
//   k_camera            sample points + camera rays for every path
//   per bounce:         k_trace (closest hit) over the live-path queue,
//                       k_shade_* (material sampling, emission, Russian
//                       roulette, next-event shadow rays) compacting survivors
//                       into the next queue with one atomic per block,
//                       [NEE] k_trace (visibility) + k_shadow_accum
//   k_film_gather       each pixel gathers its 9x9 neighbourhood of samples in
//                       the order the reference's serial splat visits them, so
//                       the film sums are the reference's, operation for
//                       operation, for equal sample radiances.
// Random numbers are addressed by (path, stream, bounce, slot) through
// Philox4x32-10, identically to the CPU oracle.
#pragma clang fp contract(off)

#include <algorithm>
#include <cstring>
#include <vector>

#include "gr_guide.h"
#include "gr_internal.h"
#include "gr_shading.h"

namespace gr {

// ---------------------------------------------------------------------------
// Host: material table and camera
// ---------------------------------------------------------------------------
namespace {

void set3(float* v, vec3 c) {
    v[0] = c.x;
    v[1] = c.y;
    v[2] = c.z;
}
vec3 lerp3h(vec3 a, vec3 b, float t) {
    return mk(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z));
}
vec3 tint(vec3 c) {  // Vec::tint, src/core/vector.rs:218-234
    const float l = 0.21f * c.x + 0.71f * c.y + 0.08f * c.z;
    return l > 0.0f ? c / l : mk(1.0f, 1.0f, 1.0f);
}
void ggx_params(GpuLobe& l, float roughness, float anisotropic) {  // util.rs:128-133
    const float aspect = sqrtf(1.0f - anisotropic * 0.9f);
    l.v[0] = fmaxf(0.001f, roughness * roughness / aspect);
    l.v[1] = fmaxf(0.001f, roughness * roughness * aspect);
}

}  // namespace

GpuMaterial build_material(const float rec[GR_MAT_FLOATS]) {
    GpuMaterial m;
    std::memset(&m, 0, sizeof(m));
    const int kind = (int)rec[0];
    const vec3 one = mk(1.0f, 1.0f, 1.0f);
    if (kind == 1) {  // Material::diffuse_light
        m.has_light = 1;
        m.light[0] = m.display[0] = rec[1];
        m.light[1] = m.display[1] = rec[2];
        m.light[2] = m.display[2] = rec[3];
        return m;
    }
    if (kind == 2 || kind == 3) {  // Material::diffuse / mirror
        m.display[0] = m.display[1] = m.display[2] = 1.0f;
        m.nlobes = 1;
        m.roughness = kind == 2 ? 1.0f : 0.0f;
        m.lobes[0].type = kind == 2 ? kLobePerfectDiffuse : kLobePerfectMirror;
        return m;
    }
    // DisneyMaterialBuilder::build, src/material/material.rs:292-333
    const vec3 base = mk(rec[1], rec[2], rec[3]);
    set3(m.display, base);  // display: the builder's base colour (material.rs:329)
    const float roughness = rec[4], anisotropic = rec[5], ior = rec[6], metallic = rec[7];
    const float specular_trans = rec[8], specular_tint = rec[9], sheen = rec[10], sheen_tint = rec[11];
    const float clearcoat = rec[12], clearcoat_gloss = rec[13];
    m.roughness = roughness;
    const float diffuse_weight = (1.0f - metallic) * (1.0f - specular_trans);
    const float trans_weight = (1.0f - metallic) * specular_trans;
    if (diffuse_weight > 0.0f) {  // DisneyDiffuseRefl::new, lobes.rs:83-91
        GpuLobe& l = m.lobes[m.nlobes++];
        l.type = kLobeDisneyDiffuse;
        set3(&l.v[0], base * diffuse_weight);
        set3(&l.v[3], lerp3h(one, tint(base), sheen_tint) * (sheen * diffuse_weight));
        l.v[6] = roughness;
    }
    if (ior > 1.0f) {  // DisneySpecularRefl::new_aniso, lobes.rs:212-225
        GpuLobe& l = m.lobes[m.nlobes++];
        l.type = kLobeDisneySpecular;
        ggx_params(l, roughness, anisotropic);
        set3(&l.v[2], base);
        set3(&l.v[5], lerp3h(one, tint(base), specular_tint));
        l.v[8] = fmaxf(ior, 1.01f);
        l.v[9] = metallic;
        set3(&l.v[10], one);
    }
    if (clearcoat > 0.0f) {  // DisneyClearcoatRefl::new, lobes.rs:232-242
        GpuLobe& l = m.lobes[m.nlobes++];
        l.type = kLobeClearcoat;
        l.v[0] = 0.1f + clearcoat_gloss * (0.001f - 0.1f);
        set3(&l.v[1], one * 0.04f);
        set3(&l.v[4], one * (0.25f * clearcoat));
    }
    if (trans_weight > 0.0f) {  // DisneySpecularTrans::new_aniso, lobes.rs:258-268
        GpuLobe& l = m.lobes[m.nlobes++];
        l.type = kLobeSpecularTrans;
        ggx_params(l, roughness, anisotropic);
        l.v[2] = fmaxf(ior, 1.01f);
        set3(&l.v[3], mk(sqrtf(base.x), sqrtf(base.y), sqrtf(base.z)) * trans_weight);
    }
    return m;
}

// Camera::new's xform = translation * rotation (src/core/camera.rs:42-61,
// src/core/matrix.rs:73-121,317-400), row-major.
void camera_matrix(const gr_camera& c, float m[16]) {
    float t[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, c.translate[0], c.translate[1], c.translate[2], 1};
    const float r0 = c.rotate[0], ix = c.rotate[1], iy = c.rotate[2], iz = c.rotate[3];
    float r[16] = {0};
    r[0] = 1.0f - 2.0f * (iy * iy + iz * iz);
    r[1] = 2.0f * (ix * iy + iz * r0);
    r[2] = 2.0f * (iz * ix - iy * r0);
    r[4] = 2.0f * (ix * iy - iz * r0);
    r[5] = 1.0f - 2.0f * (iz * iz + ix * ix);
    r[6] = 2.0f * (iy * iz + ix * r0);
    r[8] = 2.0f * (iz * ix + iy * r0);
    r[9] = 2.0f * (iy * iz - ix * r0);
    r[10] = 1.0f - 2.0f * (iy * iy + ix * ix);
    r[15] = 1.0f;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            m[4 * i + j] = t[4 * i] * r[j] + t[4 * i + 1] * r[4 + j] + t[4 * i + 2] * r[8 + j] + t[4 * i + 3] * r[12 + j];
}

// ---------------------------------------------------------------------------
// Host: shading data upload (called from gr_scene_commit)
// ---------------------------------------------------------------------------
hipError_t upload_shading(gr_scene* s, std::vector<GpuPrim>* prims_out) {
    const auto& P = s->prims;
    std::vector<GpuMaterial> mats(P.size());
    std::vector<GpuPrim> prims(P.size());
    std::vector<float4> verts, normals;
    std::vector<float2> uvs;
    std::vector<uint32_t> tris, lights;
    std::vector<float> cdf;
    for (size_t i = 0; i < P.size(); ++i) {
        const Prim& p = P[i];
        mats[i] = build_material(p.material);
        if (mats[i].has_light) lights.push_back((uint32_t)i);
        GpuPrim g;
        std::memset(&g, 0, sizeof(g));
        g.type = p.type;
        if (p.type == kPrimMesh) {
            g.tri_off = (uint32_t)tris.size();
            g.vert_off = (uint32_t)verts.size();
            g.norm_off = (uint32_t)normals.size();
            g.uv_off = (uint32_t)uvs.size();
            g.ntris = p.ntris;
            g.cdf_off = (uint32_t)cdf.size();
            for (const vec3& v : p.verts) verts.push_back(make_float4(v.x, v.y, v.z, 0.0f));
            for (const vec3& v : p.normals) normals.push_back(make_float4(v.x, v.y, v.z, 0.0f));
            for (const vec3& v : p.uvs) uvs.push_back(make_float2(v.x, v.y));
            tris.insert(tris.end(), p.tris.begin(), p.tris.end());
            // area CDF, src/geom/mesh.rs:161-170
            float total = 0.0f;
            const size_t c0 = cdf.size();
            for (uint32_t t = 0; t < p.ntris; ++t) {
                const uint32_t* tr = &p.tris[9 * (size_t)t];
                const vec3 e1 = p.verts[tr[0]] - p.verts[tr[2]];
                const vec3 e2 = p.verts[tr[1]] - p.verts[tr[2]];
                total += 0.5f * length(cross(e1, e2));
                cdf.push_back(total);
            }
            for (size_t k = c0; k < cdf.size(); ++k) cdf[k] = cdf[k] / total;
            g.area = total;
        } else {
            g.center[0] = p.center.x;
            g.center[1] = p.center.y;
            g.center[2] = p.center.z;
            g.radius = p.radius;
            g.area = 4.0f * kPi * p.radius * p.radius;
        }
        prims[i] = g;
    }
    if (prims_out) *prims_out = prims;
    if (s->ctx->device == GR_HOST_ONLY) return hipSuccess;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_mats = off; off = align(off + mats.size() * sizeof(GpuMaterial));
    const size_t o_prims = off; off = align(off + prims.size() * sizeof(GpuPrim));
    const size_t o_verts = off; off = align(off + verts.size() * sizeof(float4));
    const size_t o_norms = off; off = align(off + normals.size() * sizeof(float4));
    const size_t o_uvs = off; off = align(off + uvs.size() * sizeof(float2));
    const size_t o_tris = off; off = align(off + tris.size() * sizeof(uint32_t));
    const size_t o_cdf = off; off = align(off + cdf.size() * sizeof(float));
    const size_t o_lights = off; off = align(off + lights.size() * sizeof(uint32_t) + 4);
    char* blob = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&blob), off);
    if (e != hipSuccess) return e;
    s->d_shade_blob = blob;
    auto put = [&](size_t o, const void* src, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(blob + o, src, bytes, hipMemcpyHostToDevice, s->ctx->stream);
    };
    put(o_mats, mats.data(), mats.size() * sizeof(GpuMaterial));
    put(o_prims, prims.data(), prims.size() * sizeof(GpuPrim));
    put(o_verts, verts.data(), verts.size() * sizeof(float4));
    put(o_norms, normals.data(), normals.size() * sizeof(float4));
    put(o_uvs, uvs.data(), uvs.size() * sizeof(float2));
    put(o_tris, tris.data(), tris.size() * sizeof(uint32_t));
    put(o_cdf, cdf.data(), cdf.size() * sizeof(float));
    put(o_lights, lights.data(), lights.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    ShadeView& v = s->shade;
    v.mats = reinterpret_cast<const GpuMaterial*>(blob + o_mats);
    v.prims = reinterpret_cast<const GpuPrim*>(blob + o_prims);
    v.verts = reinterpret_cast<const float4*>(blob + o_verts);
    v.normals = reinterpret_cast<const float4*>(blob + o_norms);
    v.uvs = reinterpret_cast<const float2*>(blob + o_uvs);
    v.tris = reinterpret_cast<const uint32_t*>(blob + o_tris);
    v.cdf = reinterpret_cast<const float*>(blob + o_cdf);
    v.lights = reinterpret_cast<const uint32_t*>(blob + o_lights);
    v.num_lights = (uint32_t)lights.size();
    return e;
}

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
namespace {

enum { kStreamCamera = 0, kStreamLight = 1, kStreamFilm = 2 };
// Slots within one bounce's draws: block 0 = slots 0..3, block 1 = 4..7,
// block 2 = 8..11.
enum { kSlotLobe = 0, kSlotU1 = 1, kSlotU2 = 2, kSlotRR = 3, kSlotLight = 4, kSlotLightTri = 5,
       kSlotLightU1 = 6, kSlotLightU2 = 7, kSlotLightDir1 = 8, kSlotLightDir2 = 9,
       // guided integrator, camera stream: its Russian roulette (block 2)
       kSlotGuideRR = 10,
       // guided vertices, block 3: technique choice, mixture lobe, vMF cosine, azimuth
       kSlotGuideSel = 12, kSlotGuideLobe = 13, kSlotGuideU1 = 14, kSlotGuideU2 = 15 };

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                        uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}

// Four consecutive draws (slots 4*block .. 4*block+3) of one bounce.
__device__ __forceinline__ uint4 draws(uint64_t seed, uint64_t path, uint32_t stream, uint32_t bounce,
                                       uint32_t block) {
    return philox((uint32_t)path, (uint32_t)(path >> 32), bounce, (stream << 16) | block, (uint32_t)seed,
                  (uint32_t)(seed >> 32));
}

__device__ __forceinline__ float u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

__device__ __forceinline__ vec3 ld3(const float4 v) { return mk(v.x, v.y, v.z); }

// Mesh::compute_surface_props (src/geom/mesh.rs:190-240).
__device__ Frame mesh_frame(const ShadeView& sh, const GpuPrim& p, uint32_t comp, float u, float v, float w) {
    const uint32_t* t = sh.tris + p.tri_off + 9 * (size_t)comp;
    const vec3 a = ld3(sh.verts[p.vert_off + t[0]]), b = ld3(sh.verts[p.vert_off + t[1]]),
               c = ld3(sh.verts[p.vert_off + t[2]]);
    const vec3 e1 = a - c, e2 = b - c;
    const float2 at = sh.uvs[p.uv_off + t[6]], bt = sh.uvs[p.uv_off + t[7]], ct = sh.uvs[p.uv_off + t[8]];
    const vec3 uv1 = mk(at.x - ct.x, at.y - ct.y, 0.0f - 0.0f);
    const vec3 uv2 = mk(bt.x - ct.x, bt.y - ct.y, 0.0f - 0.0f);
    Frame f;
    f.gn = normalized(cross(e1, e2));
    if (t[3] == kNone) {
        f.n = f.gn;
    } else {
        const vec3 an = ld3(sh.normals[p.norm_off + t[3]]), bn = ld3(sh.normals[p.norm_off + t[4]]),
                   cn = ld3(sh.normals[p.norm_off + t[5]]);
        f.n = normalized((an * u + bn * v) + cn * w);
    }
    const float det = uv1.x * uv2.y + uv1.y * uv2.x;
    if (det == 0.0f) {
        dev::coord_system(f.n, &f.t, &f.b);
    } else {
        const float inv = 1.0f / det;
        const vec3 dpdu = (uv1 * uv2.y - uv2 * uv1.y) * inv;
        f.b = normalized(cross(f.n, dpdu));
        f.t = cross(f.b, f.n);
    }
    return f;
}

// Sphere::compute_surface_props (src/geom/sphere.rs:28-45).
__device__ Frame sphere_frame(const GpuPrim& p, vec3 pt) {
    Frame f;
    f.n = normalized(pt - mk(p.center[0], p.center[1], p.center[2]));
    if (nearly_zero(f.n.x) && nearly_zero(f.n.z)) f.t = mk(1.0f, 0.0f, 0.0f);
    else f.t = normalized(mk(-f.n.z, 0.0f, f.n.x));
    f.b = cross(f.n, f.t);
    f.gn = f.n;
    return f;
}

// Frame of a traversal hit: u, v are recomputed from the same stored triangle
// record and ray, so they are the values the traversal saw.
__device__ Frame hit_frame(const SceneView& sv, const ShadeView& sh, const gr_hit& h, vec3 o, vec3 d) {
    const GpuPrim& p = sh.prims[h.prim];
    if (p.type == kPrimMesh) {
        const float4* q = reinterpret_cast<const float4*>(sv.comps + h.slot);
        const float4 q0 = q[0], q1 = q[1], q2 = q[2];
        const vec3 c = mk(q0.x, q0.y, q0.z), e1 = mk(q0.w, q1.x, q1.y), e2 = mk(q1.z, q1.w, q2.x);
        const vec3 pp = cross(d, e2);
        const float inv_det = 1.0f / dot(e1, pp);
        const vec3 t = o - c;
        const float u = dot(t, pp) * inv_det;
        const vec3 qq = cross(t, e1);
        const float v = dot(d, qq) * inv_det;
        return mesh_frame(sh, p, h.comp, u, v, 1.0f - u - v);
    }
    return sphere_frame(p, o + d * h.t);
}

// Index of the first cdf entry >= target (CumulativeDistribution::ind_sample,
// src/core/random.rs:248-255).
__device__ uint32_t cdf_search(const float* cdf, uint32_t n, float target) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (cdf[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    return lo < n ? lo : n - 1;
}

struct LightPoint {
    vec3 p;
    Frame f;
    uint32_t prim;
    float point_pdf;
};

// Bvh::sample_light's point part (src/geom/bvh.rs:481-496 via Prim::sample_world,
// src/geom/mesh.rs:310-327, src/geom/sphere.rs:103-111) from the draws of
// slots 4..7.
__device__ LightPoint sample_light_point(const ShadeView& sh, uint4 r) {
    LightPoint lp;
    const uint32_t k = (uint32_t)(((uint64_t)r.x * sh.num_lights) >> 32);
    lp.prim = sh.lights[k];
    const GpuPrim& p = sh.prims[lp.prim];
    if (p.type == kPrimMesh) {
        const uint32_t ti = cdf_search(sh.cdf + p.cdf_off, p.ntris, u01(r.y));
        const uint32_t* t = sh.tris + p.tri_off + 9 * (size_t)ti;
        const vec3 a = ld3(sh.verts[p.vert_off + t[0]]), b = ld3(sh.verts[p.vert_off + t[1]]),
                   c = ld3(sh.verts[p.vert_off + t[2]]);
        const float sa = sqrtf(u01(r.z));
        const float u = 1.0f - sa, v = u01(r.w) * sa;
        const float w = 1.0f - u - v;
        lp.p = (a * u + b * v) + c * w;
        lp.f = mesh_frame(sh, p, ti, u, v, w);
        lp.point_pdf = 1.0f / p.area;
    } else {
        lp.p = mk(p.center[0], p.center[1], p.center[2]) + dev::uniform_sphere(u01(r.z), u01(r.w)) * p.radius;
        lp.f = sphere_frame(p, lp.p);
        lp.point_pdf = 1.0f / (4.0f * kPi * p.radius * p.radius);
    }
    lp.point_pdf = lp.point_pdf / (float)sh.num_lights;
    return lp;
}

__device__ __forceinline__ float light_area_pdf(const ShadeView& sh, uint32_t prim) {
    const GpuPrim& p = sh.prims[prim];
    const float pdf = p.type == kPrimMesh ? 1.0f / p.area : 1.0f / (4.0f * kPi * p.radius * p.radius);
    return pdf / (float)sh.num_lights;
}

__device__ __forceinline__ float power_heuristic(float f_pdf, float g_pdf) {  // math.rs:137-142
    const float f = 1.0f * f_pdf, g = 1.0f * g_pdf;
    return (f * f) / (f * f + g * g);
}

__device__ __forceinline__ void russian_roulette(vec3& T, uint32_t depth, float rv) {
    const float lum = dev::luminance(T);
    const float a = dev::clamp_unit(lum);
    const float live = depth > 20 ? 0.10f + a * (0.75f - 0.10f) : 0.25f + a * (1.00f - 0.25f);
    if (rv < live) T = T / live;
    else T = mk(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ uint64_t lanes_below() {
    const uint32_t lane = threadIdx.x & 63u;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Reserves n consecutive entries per thread with one atomic per block;
// returns the thread's first entry.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_reserve(uint32_t n, uint32_t* counter) {
    __shared__ uint32_t s_part[17];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, waves = (blockDim.x + 63u) >> 6;
    uint32_t incl = n;  // inclusive prefix sum over the wave
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
    }
    if (lane == 63u) s_part[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < waves; ++w) {
            const uint32_t c = s_part[w];
            s_part[w] = run;
            run += c;
        }
        s_part[16] = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    const uint32_t first = s_part[16] + s_part[wave] + incl - n;
    __syncthreads();  // s_part is reused by the next call
    return first;
}

// Appends the block's `alive` threads to a queue with one atomic per block
// (up to 16 waves): atomics on one address serialise in L2, so a queue fed
// by every block is cheaper with a quarter of the atomics of per-wave appends.
// Returns each alive thread's slot.  Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_append(bool alive, uint32_t* counter) {
    __shared__ uint32_t s_part[17];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, waves = (blockDim.x + 63u) >> 6;
    const uint64_t m = __ballot(alive);
    if (lane == 0) s_part[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < waves; ++w) {
            const uint32_t c = s_part[w];
            s_part[w] = run;
            run += c;
        }
        s_part[16] = run ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    const uint32_t slot = s_part[16] + s_part[wave] + (uint32_t)__popcll(m & lanes_below());
    __syncthreads();  // s_part is reused by the next call
    return slot;
}


// State of a guided trace (GR_INTEGRATOR_GUIDED).
constexpr uint32_t kGuideNoQuery = 0xfffffffeu;  // eligible vertex, guide not sampled yet
// vertex record: raw(8) | dir, Lb.x | Lb.yz, T lum, guide-record flag |
// primitive, lobe | bucket << 4, n c_k, lobe-record flag
constexpr uint32_t kVrecFloat4 = 5;
// learned lobe choice: incoming-angle buckets per primitive (4 |cos| bins,
// outside and inside), each with a cell for the whole primitive and one per
// octant of the shading normal (on a sphere, the part of the surface); the
// samples per lobe before a cell's statistics are used, the fixed-point scale
// and the clamp of one contribution
constexpr uint32_t kLobeBuckets = 8;
constexpr uint32_t kLobeSub = 9;
constexpr uint32_t kLobeCells = kLobeBuckets * kLobeSub;  // per primitive
constexpr uint32_t kLobeMinCount = 64;
constexpr float kLobeFixed = 65536.0f;
constexpr float kLobeYMax = 1.0e4f;

struct GuideArgs {
    float* raw;        // per query: 8 raw inputs
    float* mix;        // per query: kMixFloats
    uint32_t* gslot;   // per path: query index, kGuideNoQuery or kNone
    float4* vrec;      // per training path and depth
    float* recs;       // compacted training records
    uint32_t* rec_count;
    uint32_t rec_cap, train_stride, train_paths, drec, guided_depth, active;
    uint32_t lobes;   // GR_GUIDE_LOBES_*: the lobe picks the guide may redirect
    uint32_t select;  // GR_GUIDE_SELECT_*: how the guided integrator picks a lobe
    float select_mix; // share of the uniform choice in the lobe-choice probabilities
    float rr;         // throughput luminance below which paths play Russian roulette (0: off)
    unsigned long long* lobe_sum;  // learned lobe choice: per (prim, bucket, lobe)
    uint32_t* lobe_cnt;
    const float4* lobe_c;          // per cell: choice probabilities, x < 0 (or NaN): none yet
    const float* lobe_m;           // per cell: mean radiance a continuation brought (the lobes'
                                   // mean contribution), with lobe_c
    float alpha;
    float lo[3], inv_ext[3];
};

struct RenderArgs {
    SceneView sv;
    ShadeView sh;
    uint32_t width, height, npaths, max_depth;
    uint64_t seed, path_base;
    float cam[16];
    float wx, wy;
    RenderBuffers rb;
    GuideArgs gd;
    unsigned long long* stats;  // the context's error counters (kStat*)
    // Film gather from a ring of samples (batched jobs): iteration `it`'s samples
    // start at ring slot (ring_base + it * W * H) mod ring_cap.  ring_cap 0:
    // rb.L / rb.st hold the iterations one after another from index 0.
    uint64_t ring_base, ring_cap;
};

__device__ __forceinline__ bool finite3(vec3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

// Counts a rare event in the context's error counters.
__device__ __forceinline__ void note_stat(unsigned long long* stats, int which, bool event) {
    if (event && stats) atomicAdd(stats + which, 1ull);
}

// The conditions the reference asserts on every material sample
// (debug_assert!s, src/material/material.rs:224-226).
__device__ __forceinline__ bool bad_sample(const dev::MatSample& smp) {
    return !(finite3(smp.radiance) && isfinite(smp.pdf) && smp.pdf > 0.0f);
}

// Lobes whose reported pdf is the density their sampling draws from.  The
// clearcoat pdf lacks the cos(theta_h) Jacobian factor (src/material/util.rs:
// 266-308) and the transmission pdf uses an unnormalised half vector
// (src/material/lobes.rs:337-339); the delta mirror has no density.
__device__ __forceinline__ bool lobe_consistent(uint32_t t) {
    return t == kLobeDisneyDiffuse || t == kLobeDisneySpecular || t == kLobePerfectDiffuse;
}

// Whether a pick of a lobe of type t may be redirected by the guide, for the
// guide's lobe set (GR_GUIDE_LOBES_*): the consistent lobes, every lobe with
// a density (all but the mirror), or the diffuse lobes only.
__device__ __forceinline__ bool lobe_guided(uint32_t t, uint32_t set) {
    if (set == GR_GUIDE_LOBES_ALL) return t != kLobePerfectMirror;
    if (set == GR_GUIDE_LOBES_DIFFUSE) return t == kLobeDisneyDiffuse || t == kLobePerfectDiffuse;
    return lobe_consistent(t);
}
__device__ __forceinline__ bool lobe_guided(const GpuLobe& l, uint32_t set) { return lobe_guided(l.type, set); }

__device__ __forceinline__ uint32_t guided_lobes(const GpuMaterial& m, uint32_t set) {
    uint32_t c = 0;
    for (uint32_t k = 0; k < m.nlobes; ++k) c += lobe_guided(m.lobes[k], set) ? 1u : 0u;
    return c;
}

// Lobe-choice probabilities of the guided integrator.  The reference picks
// each of the n lobes with probability 1/n; picking lobe k with probability
// c_k > 0 instead and dividing that lobe's sample by n c_k keeps each lobe's
// contribution, and so the plain estimator's mean, for any such c.
//
// GR_GUIDE_SELECT_ALBEDO: c_k = mix / n + (1 - mix) a_k / sum a, with a_k the
// luminance of lobe k's colour times its Fresnel factor at the incoming
// angle -- the share of the energy each lobe returns -- so smooth glassy
// surfaces, where one lobe carries almost everything, spend their samples on
// it.  GR_GUIDE_SELECT_LEARNED: the same with a_k the mean contribution of
// lobe k's samples on earlier batches (k_lobe_stats, k_lobe_table).  mix > 0
// keeps every c_k positive.
__device__ __forceinline__ float lobe_albedo(const GpuLobe& l, vec3 il) {
    const float c = fabsf(il.z);
    switch (l.type) {
        case kLobeDisneyDiffuse: return dev::luminance(dev::lv3(l, 0) + dev::lv3(l, 3));
        case kLobeDisneySpecular:
        case kLobeClearcoat: return dev::luminance(mul(dev::refl_colour(l, l.type), dev::refl_fresnel(l, l.type, c)));
        case kLobeSpecularTrans: return dev::luminance(dev::lv3(l, 3)) * (1.0f - dev::dielectric(il.z, l.v[2]));
        case kLobePerfectDiffuse: return 0.5f;
        default: return 1.0f;
    }
}

struct LobeChoice {
    float c[4];
    float m;  // the cell's mean continuation radiance (lobe_m), -1 without statistics
    uint32_t pick, bucket, octant;
    // c[k] for a lane-varying k, through constant indices (keeps c in registers)
    __device__ __forceinline__ float c_at(uint32_t k) const {
        return k == 0 ? c[0] : (k == 1 ? c[1] : (k == 2 ? c[2] : c[3]));
    }
};
// Incoming-angle bucket of the learned lobe choice: 4 bins of |cos theta_i|,
// outside (il.z > 0) and inside.
__device__ __forceinline__ uint32_t lobe_bucket(vec3 il) {
    const uint32_t b = (uint32_t)fminf(fabsf(il.z) * 4.0f, 3.0f);
    return il.z > 0.0f ? b : 4u + b;
}
__device__ __forceinline__ uint32_t normal_octant(vec3 n) {
    return (n.x > 0.0f ? 1u : 0u) | (n.y > 0.0f ? 2u : 0u) | (n.z > 0.0f ? 4u : 0u);
}
__device__ __forceinline__ LobeChoice choose_lobe(const GpuMaterial& m, vec3 il, vec3 n_world, const GuideArgs& g,
                                                  uint32_t u, uint32_t prim) {
    LobeChoice ch;
    const uint32_t n = m.nlobes;
    for (uint32_t k = 0; k < 4; ++k) ch.c[k] = k < n ? 1.0f / (float)n : 0.0f;
    ch.pick = (uint32_t)(((uint64_t)u * n) >> 32);  // the reference's choice
    ch.bucket = lobe_bucket(il);
    ch.octant = normal_octant(n_world);
    ch.m = -1.0f;
    if (g.select == GR_GUIDE_SELECT_UNIFORM || n < 2) return ch;
    bool learned = false;
    if (g.select == GR_GUIDE_SELECT_LEARNED) {
        const size_t cell = ((size_t)prim * kLobeBuckets + ch.bucket) * kLobeSub + 1 + ch.octant;
        const float4 c = g.lobe_c[cell];
        if (c.x >= 0.0f) {
            ch.c[0] = c.x, ch.c[1] = c.y, ch.c[2] = c.z, ch.c[3] = c.w;
            ch.m = g.lobe_m[cell];
            learned = true;
        }
    }
    // loops over the 4 lobe slots with constant indices, so a and c stay in registers
    if (!learned) {
        float a[4] = {0.0f, 0.0f, 0.0f, 0.0f}, sum = 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
            if (k < n) {
                a[k] = fmaxf(lobe_albedo(m.lobes[k], il), 0.0f);
                sum += a[k];
            }
        if (!(sum > 0.0f) || !isfinite(sum)) return ch;
        const float floor_c = g.select_mix / (float)n;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
            if (k < n) ch.c[k] = floor_c + (1.0f - g.select_mix) * (a[k] / sum);
    }
    const float x = u01(u);
    float acc = 0.0f;
    ch.pick = n - 1;
    bool found = false;
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k)
        if (!found && k + 1 < n) {
            acc += ch.c[k];
            if (x < acc) {
                ch.pick = k;
                found = true;
            }
        }
    return ch;
}

// The guided integrator's vertex at direction wo, in one pass over the
// material's lobes.
//
// sample_world's draw of lobe k at wo reports radiance F_k(wo) and pdf P(wo):
// P is the mean of the n reported lobe pdfs, and F_k is lobe k's f plus, when
// k's kind matches the reflect/transmit side of wo, every other lobe's f
// (src/material/material.rs:177-211).  A choice of lobe k with probability
// c_k weighs F_k cos / (P n c_k).
//
// With s_k the density lobe k's draws realise (lobe_density), the picks in
// the guided set G contribute (1/n) sum_{k in G} int s_k F_k cos / P dwo.  A
// guided vertex enters G with its lobe choice (probability c_G, the sum of
// c_k over G) and draws wo from D = alpha q + (1 - alpha) sum_{k in G}
// (c_k / c_G) s_k (q: the guide's mixture), so the weight
//   sum_{k in G} s_k F_k cos / (n P c_G D)
// has the same mean.
struct ChoiceEval {
    vec3 f_sum, f_pick;  // all lobes' f; the chosen lobe's
    vec3 own;            // sum over G of s_k f_k where F_k = f_k
    float s_all;         // sum of the reported pdfs
    float s_take;        // sum over G of s_k where F_k = f_sum
    float c_g, cs_g;     // sums over G of c_k and c_k s_k
    bool take_pick;      // F_pick = f_sum
};
__device__ __forceinline__ ChoiceEval choice_eval(const GpuMaterial& m, vec3 il, vec3 ol, bool refl, uint32_t set,
                                                  const float* c, uint32_t pick, bool in_g) {
    ChoiceEval e;
    e.f_sum = e.f_pick = e.own = mk(0.0f, 0.0f, 0.0f);
    e.s_all = e.s_take = e.c_g = e.cs_g = 0.0f;
    e.take_pick = false;
    // lobe by lobe in position order (a walk by type: the lane's c_k selected
    // with constant indices, which keeps c in registers)
    dev::each_lobe(m, [&](uint32_t k, const GpuLobe& l, uint32_t t) {
        vec3 fk;
        float pk;
        dev::lobe_f_pdf(l, t, il, ol, true, true, &fk, &pk);
        e.f_sum = e.f_sum + fk;
        if (il.z != 0.0f) e.s_all += pk;
        const uint32_t kind = dev::lobe_kind(t);
        const bool take = (refl && (kind & kLobeReflection)) || (!refl && (kind & kLobeTransmission));
        if (k == pick) {
            e.f_pick = fk;
            e.take_pick = take;
        }
        if (!in_g || !lobe_guided(t, set)) return;
        const float sk = (t == kLobeClearcoat || t == kLobeSpecularTrans) ? dev::lobe_density(l, t, il, ol) : pk;
        const float ck = k == 0 ? c[0] : (k == 1 ? c[1] : (k == 2 ? c[2] : c[3]));
        e.c_g += ck;
        e.cs_g += ck * sk;
        if (take)
            e.s_take += sk;
        else
            e.own = e.own + fk * sk;
    });
    return e;
}

// Vertex-record slot of path p at this depth, or kNone when p records no
// training data there.
__device__ __forceinline__ uint32_t vrec_slot(const GuideArgs& g, uint32_t p, uint32_t depth) {
    if (g.train_stride == 0 || p % g.train_stride != 0 || depth >= g.drec) return kNone;
    const uint32_t r = p / g.train_stride;
    return r < g.train_paths ? r * g.drec + depth : kNone;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

struct CameraSample {
    float2 st;
    vec3 o, d;
    float cos_t;  // cosine between the camera-space direction and the view axis (-z)
};

// Sample point (film.rs:55-76) and camera ray (camera.rs:87-97) of one path.
__device__ __forceinline__ CameraSample camera_sample(const float* m, float wx, float wy, uint32_t width,
                                                      uint32_t height, uint64_t seed, uint64_t path, uint32_t row,
                                                      uint32_t col) {
    CameraSample c;
    const uint4 r = draws(seed, path, kStreamFilm, 0, 0);
    const float rj = (0.5f + (float)row) + (-2.0f + 2.0f * 2.0f * u01(r.x));
    const float cj = (0.5f + (float)col) + (-2.0f + 2.0f * 2.0f * u01(r.y));
    const float s = dev::lerp(-1.0f, 1.0f, cj / (float)width);
    const float t = dev::lerp(-1.0f, 1.0f, rj / (float)height);
    c.st = make_float2(s, t);
    const vec3 dir = normalized(mk(wx * s, wy * t, -1.0f));
    const float z = 0.0f;
    const float x = z * m[0] + z * m[4] + z * m[8] + m[12];
    const float y = z * m[1] + z * m[5] + z * m[9] + m[13];
    const float zz = z * m[2] + z * m[6] + z * m[10] + m[14];
    const float w = z * m[3] + z * m[7] + z * m[11] + m[15];
    c.o = mk(x / w, y / w, zz / w);
    c.d = mk(dir.x * m[0] + dir.y * m[4] + dir.z * m[8], dir.x * m[1] + dir.y * m[5] + dir.z * m[9],
             dir.x * m[2] + dir.y * m[6] + dir.z * m[10]);
    c.cos_t = -dir.z;
    return c;
}

// Solid-angle density of one camera ray through a window point whose
// camera-space direction makes cos_t with the view axis: sample points cover
// the film with density 1 per pixel area and iteration (a pixel's point is
// uniform over the 4x4 pixels around it), i.e. 1 / (W H) per path over the
// window of area (2 wx)(2 wy) at distance 1, whose solid-angle element is
// cos^3 theta dA.  BDPT's light-tracing strategy and its MIS weights use it.
__device__ __forceinline__ float camera_dir_pdf(float wx, float wy, float cos_t) {
    return 1.0f / (((4.0f * wx) * wy) * ((cos_t * cos_t) * cos_t));
}

__global__ __launch_bounds__(256) void k_camera(RenderArgs a) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npaths) return;
    const uint32_t npx = a.width * a.height;
    const uint32_t i = p % npx;
    const CameraSample c =
        camera_sample(a.cam, a.wx, a.wy, a.width, a.height, a.seed, a.path_base + p, i / a.width, i % a.width);
    a.rb.st[p] = c.st;
    a.rb.rays[0][2 * (size_t)p] = make_float4(c.o.x, c.o.y, c.o.z, 0.0f);
    a.rb.rays[0][2 * (size_t)p + 1] = make_float4(c.d.x, c.d.y, c.d.z, 0.0f);
    a.rb.queue[0][p] = p;
    a.rb.T[p] = make_float4(1.0f, 1.0f, 1.0f, camera_dir_pdf(a.wx, a.wy, c.cos_t));  // w: BDPT's ray density
    a.rb.L[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    a.rb.aux[p] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);  // previous bsdf pdf, previous bounce specular
    if (p == 0) {
        a.rb.counts[0] = a.npaths;
        a.rb.counts[1] = 0;
        a.rb.counts[2] = 0;
    }
}

__global__ void k_reset(uint32_t* c, uint32_t which) { c[which] = 0u; }

// DisplayColorIntegrator::integrate (src/render/integrators.rs:21-34): the
// display colour of the camera ray's closest hit, black on a miss.
__global__ __launch_bounds__(256) void k_display(RenderArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.npaths) return;
    const uint32_t p = a.rb.queue[0][i];
    const uint32_t prim = a.rb.hits[i].x;
    float4 L = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (prim != kNone) {
        const float* c = a.sh.mats[prim].display;
        L = make_float4(c[0], c[1], c[2], 0.0f);
    }
    a.rb.L[p] = L;
}

// One bounce of PathTracerIntegrator::integrate (src/render/integrators.rs:42-99);
// with NEE, light sampling and power-heuristic MIS.
//
// With GUIDED, the reference's BSDF estimator -- pick one of the n lobes
// uniformly, sample it, weight F / ((1/n) sum_k pdf_k) -- is split by the
// picked lobe.  A pick in the guide's lobe set G (lobe_guided) takes its
// direction from the guide's mixture q with probability alpha and from the
// picked lobe otherwise, weighted as choice_eval derives so that the picks
// in G keep the mean they have in the reference's estimator; any other pick
// is the reference's sample as is.  So the guided estimator converges to the
// plain estimator's image, including the reference's pdf quirks; the NEE and
// emission MIS weights use the reported BSDF pdf exactly as without the guide.
struct ShadeOut {
    bool alive, shadow;
    vec3 new_o, new_d, seg_to, c_nee, hp;
};

// One bounce of path `path` (state slot p: T, L, aux) at hit h of ray (o, d)
// in scene (sv, sh).  Shared by the per-film wavefront and the batch queue so
// both compute the same arithmetic.
template <bool NEE, bool GUIDED, bool CHOICE>
__device__ __forceinline__ void shade_path(const RenderArgs& a, const SceneView& sv, const ShadeView& sh,
                                           uint64_t seed, uint32_t max_depth, uint64_t path, uint32_t p,
                                           uint32_t depth, const gr_hit& h, vec3 o, vec3 d, ShadeOut& out) {
    bool alive = false;
    bool shadow = false;
    vec3 new_o = mk(0.0f, 0.0f, 0.0f), new_d = new_o, seg_to = new_o, c_nee = new_o, hp = new_o;
    {
        vec3 T = ld3(a.rb.T[p]);
        vec3 L = ld3(a.rb.L[p]);
        if (h.prim != kNone) {
            const Frame f = hit_frame(sv, sh, h, o, d);
            const GpuMaterial& m = sh.mats[h.prim];
            const vec3 wi = -d;
            hp = o + d * h.t;
            const uint4 r0 = draws(seed, path, kStreamCamera, depth, 0);
            if (!NEE) {
                const dev::MatSample smp = dev::sample_world(m, wi, f, true, r0.x, u01(r0.y), u01(r0.z));
                note_stat(a.stats, kStatNonfiniteBsdf, bad_sample(smp));
                L = L + mul(T, smp.emission);
                T = mul(T, smp.radiance * (fabsf(dot(f.n, smp.outgoing)) / smp.pdf));
                new_o = hp + smp.outgoing * kRayPush;
                new_d = smp.outgoing;
                if (depth > 10 || dev::nearly_zero3(T)) russian_roulette(T, depth, u01(r0.w));
                alive = !dev::is_zero(T);
            } else {
                const float4 aux = a.rb.aux[p];
                const vec3 le = dev::light_world(m, wi, f);
                if (!dev::is_zero(le)) {
                    if (aux.y != 0.0f) {
                        L = L + mul(T, le);
                    } else {
                        const float pl = light_area_pdf(sh, h.prim) * (h.t * h.t) / fabsf(dot(f.gn, wi));
                        L = L + mul(T, le) * power_heuristic(aux.x, pl);
                    }
                }
                uint32_t gs = kNone;
                const float* mix = nullptr;
                if (GUIDED) {
                    gs = a.gd.gslot[p];
                    if (gs < kGuideNoQuery) mix = a.gd.mix + (size_t)gs * kMixFloats;
                }
                const float alpha = mix ? a.gd.alpha : 0.0f;
                if (m.nlobes != 0) {
                    if (sh.num_lights > 0 && dev::count_lobes(m, kLobeDiffuse | kLobeGlossy) != 0) {
                        const LightPoint lp = sample_light_point(sh, draws(seed, path, kStreamCamera, depth, 1));
                        const vec3 to = lp.p - hp;
                        const float d2 = dot(to, to);
                        const float dist = sqrtf(d2);
                        const vec3 wo = to / dist;
                        const vec3 le_l = dev::light_world(sh.mats[lp.prim], -wo, lp.f);
                        const float cos_l = fabsf(dot(lp.f.gn, wo));
                        if (!dev::is_zero(le_l) && cos_l > 0.0f && d2 > 0.0f) {
                            // f_world and pdf_world in one pass over the lobes (the same
                            // values; the reverse density it could give is not used)
                            const dev::FPdf fp = dev::f_pdf_world(m, wi, wo, f, true);
                            const vec3 fr = fp.f;
                            if (!dev::is_zero(fr)) {
                                const float pdf_light = lp.point_pdf * d2 / cos_l;
                                const float pdf_bsdf = fp.pdf_fwd;
                                const float w = power_heuristic(pdf_light, pdf_bsdf);
                                c_nee = mul(mul(T, fr), le_l) * (fabsf(dot(f.n, wo)) * w / pdf_light);
                                seg_to = lp.p;
                                shadow = true;
                            }
                        }
                    }
                    const vec3 T_in = T;
                    vec3 out_dir;
                    // the lobe the reference's sample_world picks with this draw, or the
                    // guided integrator's lobe choice (choose_lobe)
                    LobeChoice ch{};
                    if (CHOICE) {
                        ch = choose_lobe(m, dev::to_local(wi, f), f.n, a.gd, r0.x, h.prim);
                    } else {
                        ch.pick = (uint32_t)(((uint64_t)r0.x * m.nlobes) >> 32);
                    }
                    const uint32_t pick = ch.pick;
                    // the pixel's scale for the guided roulette: the first vertex's
                    // mean continuation radiance, carried in aux.z
                    const float m_first = !CHOICE ? 0.0f : (depth == 0 ? fmaxf(ch.m, 0.0f) : aux.z);
                    const bool split = GUIDED && gs != kNone && lobe_guided(m.lobes[pick], a.gd.lobes);
                    if (!CHOICE) {
                        const dev::MatSample smp = dev::sample_world(m, wi, f, true, r0.x, u01(r0.y), u01(r0.z));
                        note_stat(a.stats, kStatNonfiniteBsdf, bad_sample(smp));
                        T = mul(T, smp.radiance * (fabsf(dot(f.n, smp.outgoing)) / smp.pdf));
                        a.rb.aux[p] = make_float4(smp.pdf, (smp.kind & kLobeSpecular) ? 1.0f : 0.0f, 0.0f, 0.0f);
                        out_dir = smp.outgoing;
                    } else {
                        // the direction: the guide's mixture (a guided pick, probability
                        // alpha) or the chosen lobe's own draw; then one pass over the lobes
                        const vec3 il = dev::to_local(wi, f);
                        bool from_guide = false;
                        uint4 rg = make_uint4(0u, 0u, 0u, 0u);
                        if (split) {
                            rg = draws(seed, path, kStreamCamera, depth, kSlotGuideSel / 4);
                            from_guide = mix && u01(rg.x) < alpha;
                        }
                        if (from_guide) {
                            out_dir = mixture_sample(mix, u01(rg.y), u01(rg.z), u01(rg.w));
                        } else {
                            vec3 hh = mk(0.0f, 0.0f, 0.0f);
                            out_dir = dev::to_world(dev::lobe_draw(m.lobes[pick], il, u01(r0.y), u01(r0.z), &hh), f);
                        }
                        const GpuLobe& lp = m.lobes[pick];
                        const float nc = (float)m.nlobes * ch.c_at(pick);
                        vec3 g = mk(0.0f, 0.0f, 0.0f);
                        float pdf_rep = 1.0f;
                        bool spec = false;
                        if (!dev::is_zero(out_dir)) {
                            const vec3 ol = dev::to_local(out_dir, f);
                            const float cos_o = fabsf(dot(f.n, out_dir));
                            if (dev::lobe_kind(lp) & kLobeSpecular) {
                                // a delta lobe: its own f and pdf, no other lobe (sample_world)
                                pdf_rep = 1.0f / (float)m.nlobes;
                                g = dev::lobe_f(lp, il, ol, true) * (cos_o / pdf_rep / nc);
                                spec = true;
                            } else {
                                const bool refl = (dot(wi, f.gn) * dot(out_dir, f.gn)) > 0.0f;
                                const ChoiceEval ev = choice_eval(m, il, ol, refl, a.gd.lobes, ch.c, pick, split);
                                pdf_rep = ev.s_all / (float)m.nlobes;
                                if (pdf_rep > 0.0f) {
                                    if (!split) {
                                        g = (ev.take_pick ? ev.f_sum : ev.f_pick) * (cos_o / pdf_rep / nc);
                                    } else {
                                        const float q = mix ? mixture_pdf(mix, out_dir) : 0.0f;
                                        const float dens = alpha * q + (1.0f - alpha) * (ev.cs_g / ev.c_g);
                                        if (dens > 0.0f)
                                            g = (ev.f_sum * ev.s_take + ev.own) *
                                                (cos_o / (((float)m.nlobes * pdf_rep) * (ev.c_g * dens)));
                                    }
                                } else {
                                    pdf_rep = 1.0f;  // sample_world's black sample
                                }
                            }
                        }
                        note_stat(a.stats, kStatNonfiniteBsdf, !finite3(g));
                        T = mul(T, g);
                        a.rb.aux[p] = make_float4(pdf_rep, spec ? 1.0f : 0.0f, m_first, 0.0f);
                    }
                    // training records: the direction and the radiance so far (this
                    // vertex's NEE is added when its shadow ray resolves) for the guide
                    // at split vertices, and for the learned lobe choice the primitive,
                    // incoming bucket, lobe and n c_k at every vertex with a choice
                    const bool lobe_rec = CHOICE && a.gd.select == GR_GUIDE_SELECT_LEARNED && m.nlobes > 1;
                    if (split || lobe_rec) {
                        const uint32_t vs = vrec_slot(a.gd, p, depth);
                        if (vs != kNone) {
                            float4* vr = a.gd.vrec + (size_t)vs * kVrecFloat4;
                            vr[2] = make_float4(out_dir.x, out_dir.y, out_dir.z, L.x);
                            vr[3] = make_float4(L.y, L.z, dev::luminance(T_in), split ? 1.0f : 0.0f);
                            vr[4] = make_float4(__uint_as_float(h.prim),
                                                __uint_as_float(pick | (ch.bucket << 4) | (ch.octant << 8)),
                                                (float)m.nlobes * ch.c_at(pick), lobe_rec ? 1.0f : 0.0f);
                        }
                    }
                    new_o = hp + out_dir * kRayPush;
                    new_d = out_dir;
                    if (CHOICE && a.gd.rr > 0.0f && depth >= 1) {
                        // the guided integrator's roulette: a continuation whose expected
                        // share of its pixel -- throughput times this vertex's mean
                        // continuation radiance over the first vertex's -- is below rr
                        // goes on with probability share / rr, its throughput divided by
                        // that (unbiased); without those statistics the share is the
                        // throughput's luminance
                        const float share = (m_first > 0.0f && ch.m >= 0.0f)
                                                ? dev::luminance(T_in) * (ch.m / m_first)
                                                : dev::luminance(T);
                        if (share < a.gd.rr) {
                            const float q = share / a.gd.rr;
                            const uint4 rq = draws(seed, path, kStreamCamera, depth, kSlotGuideRR / 4);
                            T = (u01(rq.z) < q) ? T / q : mk(0.0f, 0.0f, 0.0f);
                        }
                    }
                    if (depth > 10 || dev::nearly_zero3(T)) russian_roulette(T, depth, u01(r0.w));
                    alive = !dev::is_zero(T);
                }
            }
        }
        a.rb.T[p] = make_float4(T.x, T.y, T.z, 0.0f);
        a.rb.L[p] = make_float4(L.x, L.y, L.z, 0.0f);
        if (max_depth != 0 && depth + 1 >= max_depth) alive = false;
    }
    out.alive = alive;
    out.shadow = shadow;
    out.new_o = new_o;
    out.new_d = new_d;
    out.seg_to = seg_to;
    out.c_nee = c_nee;
    out.hp = hp;
}

// GUIDED: the guide's mixture may be sampled at this depth (the guide's
// queries ran); CHOICE: the guided integrator's lobe choice and records.
template <bool NEE, bool GUIDED, bool CHOICE>
__global__ __launch_bounds__(256) void k_shade(RenderArgs a, uint32_t depth, uint32_t cur) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = a.rb.counts[cur];
    ShadeOut out;
    out.alive = out.shadow = false;
    uint32_t p = 0;
    if (i < n) {
        p = a.rb.queue[cur][i];
        const uint4 hv = a.rb.hits[i];
        const gr_hit h = {hv.x, hv.y, __uint_as_float(hv.z), hv.w};
        const vec3 o = ld3(a.rb.rays[cur][2 * (size_t)i]), d = ld3(a.rb.rays[cur][2 * (size_t)i + 1]);
        shade_path<NEE, GUIDED, CHOICE>(a, a.sv, a.sh, a.seed, a.max_depth, a.path_base + p, p, depth, h, o, d, out);
    }
    const uint32_t nxt = cur ^ 1u;
    const uint32_t slot = block_append(out.alive, &a.rb.counts[nxt]);
    if (out.alive) {
        a.rb.queue[nxt][slot] = p;
        a.rb.rays[nxt][2 * (size_t)slot] = make_float4(out.new_o.x, out.new_o.y, out.new_o.z, 0.0f);
        a.rb.rays[nxt][2 * (size_t)slot + 1] = make_float4(out.new_d.x, out.new_d.y, out.new_d.z, 0.0f);
    }
    if (NEE) {
        const uint32_t ss = block_append(out.shadow, &a.rb.counts[2]);
        if (out.shadow) {
            a.rb.seg[2 * (size_t)ss] = make_float4(out.hp.x, out.hp.y, out.hp.z, 0.0f);
            a.rb.seg[2 * (size_t)ss + 1] = make_float4(out.seg_to.x, out.seg_to.y, out.seg_to.z, 0.0f);
            a.rb.contrib[ss] = make_float4(out.c_nee.x, out.c_nee.y, out.c_nee.z, 0.0f);
            a.rb.seg_path[ss] = p;
        }
    }
}

template <bool GUIDED>
__global__ __launch_bounds__(256) void k_shadow_accum(RenderArgs a, uint32_t depth) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.rb.counts[2] || a.rb.vis[j] == 0) return;
    const uint32_t p = a.rb.seg_path[j];
    const float4 c = a.rb.contrib[j];
    float4 L = a.rb.L[p];
    L.x = L.x + c.x;
    L.y = L.y + c.y;
    L.z = L.z + c.z;
    a.rb.L[p] = L;
    if (GUIDED) {
        // this vertex's own light sample is not part of the radiance its
        // sampled direction brings in
        const uint32_t vs = vrec_slot(a.gd, p, depth);
        if (vs != kNone) {
            float4* vr = a.gd.vrec + (size_t)vs * kVrecFloat4;
            if (vr[3].w != 0.0f || vr[4].w != 0.0f) {
                vr[2].w = vr[2].w + c.x;
                vr[3].x = vr[3].x + c.y;
                vr[3].y = vr[3].y + c.z;
            }
        }
    }
}

// Guide inputs of the vertices the guide covers (depth < guided_depth, a lobe
// choice in the guided set): queries when the guide is sampled, vertex
// records for training paths.
__global__ __launch_bounds__(256) void k_guide_query(RenderArgs a, uint32_t depth, uint32_t cur) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = a.rb.counts[cur];
    bool eligible = false;
    uint32_t p = 0;
    float x[8];
    if (i < n) {
        p = a.rb.queue[cur][i];
        const uint4 hv = a.rb.hits[i];
        const gr_hit h = {hv.x, hv.y, __uint_as_float(hv.z), hv.w};
        if (h.prim != kNone && (a.gd.guided_depth == 0 || depth < a.gd.guided_depth)) {
            const GpuMaterial& m = a.sh.mats[h.prim];
            if (guided_lobes(m, a.gd.lobes) != 0) {
                const vec3 o = ld3(a.rb.rays[cur][2 * (size_t)i]), d = ld3(a.rb.rays[cur][2 * (size_t)i + 1]);
                const Frame f = hit_frame(a.sv, a.sh, h, o, d);
                // the lobe choice the shading kernel makes with the same draw and
                // table: only a pick in the guided set uses the mixture
                const uint4 r0 = draws(a.seed, a.path_base + p, kStreamCamera, depth, 0);
                const LobeChoice ch = choose_lobe(m, dev::to_local(-d, f), f.n, a.gd, r0.x, h.prim);
                eligible = lobe_guided(m.lobes[ch.pick], a.gd.lobes);
                if (eligible) guide_raw_inputs(o + d * h.t, -d, f.n, m.roughness, a.gd.lo, a.gd.inv_ext, x);
            }
        }
    }
    const bool query = eligible && a.gd.active;
    const uint32_t qi = block_append(query, &a.rb.counts[3]);
    if (i < n) a.gd.gslot[p] = query ? qi : (eligible ? kGuideNoQuery : kNone);
    if (query) {
        float4* r = reinterpret_cast<float4*>(a.gd.raw + (size_t)qi * 8);
        r[0] = make_float4(x[0], x[1], x[2], x[3]);
        r[1] = make_float4(x[4], x[5], x[6], x[7]);
    }
    if (eligible) {
        const uint32_t vs = vrec_slot(a.gd, p, depth);
        if (vs != kNone) {
            float4* vr = a.gd.vrec + (size_t)vs * kVrecFloat4;
            vr[0] = make_float4(x[0], x[1], x[2], x[3]);
            vr[1] = make_float4(x[4], x[5], x[6], x[7]);
        }
    }
}

// Vertex record s -> training record (raw, direction, f cos Li / pdf as
// luminance: the radiance the path gathered after the vertex, divided by the
// throughput it arrived with).  *had: the slot held a record.
__device__ __forceinline__ bool guide_record(const RenderArgs& a, uint32_t s, float4* r0, float4* r1, float4* r2,
                                             float* v, bool* had) {
    const float4* vr = a.gd.vrec + (size_t)s * kVrecFloat4;
    const float4 r3 = vr[3];
    *had = r3.w != 0.0f;
    if (!*had) return false;
    *r0 = vr[0];
    *r1 = vr[1];
    *r2 = vr[2];
    const uint32_t p = (s / a.gd.drec) * a.gd.train_stride;
    const float4 L = p < a.npaths ? a.rb.L[p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const vec3 dl = mk(L.x - r2->w, L.y - r3.x, L.z - r3.y);
    const float tl = r3.z;
    *v = tl > 0.0f ? fmaxf(dev::luminance(dl), 0.0f) / tl : 0.0f;
    return (*v == *v) && *v < 3.0e38f;
}

// Training records are compacted in slot order (count, scan, write), so the
// record list -- and the training batches drawn from it -- do not depend on
// the order in which blocks run.
__global__ __launch_bounds__(256) void k_guide_rec_count(RenderArgs a, uint32_t nslots, uint32_t* bcount) {
    __shared__ uint32_t s_c[4];
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    float4 r0, r1, r2;
    float v = 0.0f;
    bool had = false;
    const bool ok = s < nslots && guide_record(a, s, &r0, &r1, &r2, &v, &had);
    const uint64_t m = __ballot(ok);
    if ((threadIdx.x & 63u) == 0) s_c[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
}

// Exclusive scan of the per-block counts (one block); *total = their sum.
__global__ __launch_bounds__(1024) void k_guide_rec_scan(const uint32_t* bcount, uint32_t nblocks, uint32_t* boff,
                                                         uint32_t* total) {
    __shared__ uint32_t s_sum[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nblocks + 1023u) / 1024u;
    const uint32_t b0 = t * per, b1 = b0 + per < nblocks ? b0 + per : nblocks;
    uint32_t run = 0;
    for (uint32_t b = b0; b < b1; ++b) run += bcount[b];
    s_sum[t] = run;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t add = t >= off ? s_sum[t - off] : 0u;
        __syncthreads();
        s_sum[t] += add;
        __syncthreads();
    }
    run = t ? s_sum[t - 1] : 0u;
    for (uint32_t b = b0; b < b1; ++b) {
        boff[b] = run;
        run += bcount[b];
    }
    if (t == 1023u) *total = s_sum[1023];
}

// Writes each record at its block's offset plus its rank in the block, and
// clears the vertex records for the next batch.
__global__ __launch_bounds__(256) void k_guide_rec_write(RenderArgs a, uint32_t nslots, const uint32_t* boff) {
    __shared__ uint32_t s_c[4];
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    float4 r0, r1, r2;
    float v = 0.0f;
    bool had = false;
    const bool ok = s < nslots && guide_record(a, s, &r0, &r1, &r2, &v, &had);
    const uint64_t m = __ballot(ok);
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) s_c[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t k = boff[blockIdx.x] + (uint32_t)__popcll(m & lanes_below());
    for (uint32_t w = 0; w < wave; ++w) k += s_c[w];
    if (ok && k < a.gd.rec_cap) {
        float* o = a.gd.recs + (size_t)k * kRecordFloats;
        o[0] = r0.x, o[1] = r0.y, o[2] = r0.z, o[3] = r0.w;
        o[4] = r1.x, o[5] = r1.y, o[6] = r1.z, o[7] = r1.w;
        o[8] = r2.x, o[9] = r2.y, o[10] = r2.z, o[11] = v;
    }
    if (had) a.gd.vrec[(size_t)s * kVrecFloat4 + 3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Learned lobe choice, after each traced batch: every vertex record with a
// lobe choice adds |Y| n c_k -- the radiance the chosen lobe's sample brought
// (the path's radiance gathered after the vertex over the throughput it
// arrived with), without the choice's 1 / (n c_k) -- to its (primitive,
// incoming bucket, lobe) cell, in 1/65536 fixed point so the sums do not
// depend on the order of the additions; then clears the record.
__global__ __launch_bounds__(256) void k_lobe_stats(RenderArgs a, uint32_t nslots, uint32_t nprims) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    float4* vr = a.gd.vrec + (size_t)s * kVrecFloat4;
    const float4 r4 = vr[4];
    if (r4.w == 0.0f) return;
    vr[4] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 r2 = vr[2], r3 = vr[3];
    const uint32_t p = (s / a.gd.drec) * a.gd.train_stride;
    const float4 L = p < a.npaths ? a.rb.L[p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const vec3 dl = mk(L.x - r2.w, L.y - r3.x, L.z - r3.y);
    float y = r3.z > 0.0f ? fmaxf(dev::luminance(dl), 0.0f) / r3.z * r4.z : 0.0f;
    if (!(y >= 0.0f)) y = 0.0f;
    y = fminf(y, kLobeYMax);
    const uint32_t prim = __float_as_uint(r4.x), code = __float_as_uint(r4.y);
    const uint32_t lobe = code & 15u, bucket = (code >> 4) & 15u, oct = code >> 8;
    if (prim >= nprims || bucket >= kLobeBuckets || lobe >= 4u || oct >= 8u) return;  // not this scene's
    // the normal octant's cell (its primitive's cell is the sum of its 8 octants, k_lobe_table)
    const size_t cell = ((size_t)prim * kLobeBuckets + bucket) * kLobeSub + 1 + oct;
    atomicAdd(a.gd.lobe_sum + cell * 4 + lobe, (unsigned long long)(y * kLobeFixed));
    atomicAdd(a.gd.lobe_cnt + cell * 4 + lobe, 1u);
}

// Choice probabilities per cell from the sums: c_k = mix / n + (1 - mix) m_k /
// sum m with m_k the mean of lobe k's contributions, from the normal
// octant's own sums once each lobe there has kLobeMinCount samples, else from
// its primitive's (the sums of its 8 octants) once that has them (x = -1
// until then: choose_lobe falls back to the albedo rule).
__device__ __forceinline__ bool lobe_probs(const unsigned long long* sum, const uint32_t* cnt, uint32_t n,
                                           float mix, float4* out, float* m_out) {
    float mk_[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tot = 0.0f;
    for (uint32_t k = 0; k < n; ++k) {
        if (cnt[k] < kLobeMinCount) return false;
        mk_[k] = (float)((double)sum[k] / (double)cnt[k]);
        tot += mk_[k];
    }
    if (!(tot > 0.0f)) return false;
    float cc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < n; ++k) cc[k] = mix / (float)n + (1.0f - mix) * (mk_[k] / tot);
    *out = make_float4(cc[0], cc[1], cc[2], cc[3]);
    *m_out = tot / ((float)n * kLobeFixed);
    return true;
}
__global__ __launch_bounds__(256) void k_lobe_table(const unsigned long long* sum, const uint32_t* cnt,
                                                    const GpuMaterial* mats, uint32_t ncells, float mix,
                                                    float4* c, float* m) {
    const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncells) return;
    const uint32_t n = mats[cell / kLobeCells].nlobes;
    const uint32_t sub = cell % kLobeSub;
    float4 out = make_float4(-1.0f, 0.0f, 0.0f, 0.0f);
    float mo = -1.0f;
    if (sub != 0 && n >= 2 && n <= 4 && !lobe_probs(sum + (size_t)cell * 4, cnt + (size_t)cell * 4, n, mix, &out, &mo)) {
        unsigned long long ps[4] = {0ull, 0ull, 0ull, 0ull};
        uint32_t pc[4] = {0u, 0u, 0u, 0u};
        const size_t first = (size_t)(cell - sub + 1) * 4;
        for (uint32_t o = 0; o < 8; ++o)
            for (uint32_t k = 0; k < n; ++k) {
                ps[k] += sum[first + o * 4 + k];
                pc[k] += cnt[first + o * 4 + k];
            }
        (void)lobe_probs(ps, pc, n, mix, &out, &mo);
    }
    c[cell] = out;
    m[cell] = mo;
}

__global__ __launch_bounds__(256) void k_export_samples(RenderArgs a, float4* out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < a.npaths) out[p] = a.rb.L[p];
}

// mitchell_filter1 with B = C = 1/3 (src/core/math.rs:92-108).
__device__ __forceinline__ float mitchell1(float x) {
    const float B = 1.0f / 3.0f, C = 1.0f / 3.0f;
    const float tx = fabsf(2.0f * x);
    if (tx > 1.0f)
        return ((-B - 6.0f * C) * (tx * tx * tx) + (6.0f * B + 30.0f * C) * (tx * tx) + (-12.0f * B - 48.0f * C) * tx +
                (8.0f * B + 24.0f * C)) * (1.0f / 6.0f);
    return ((12.0f - 9.0f * B - 6.0f * C) * (tx * tx * tx) + (-18.0f + 12.0f * B + 6.0f * C) * (tx * tx) +
            (6.0f - 2.0f * B)) * (1.0f / 6.0f);
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ int f2i(float f) {  // saturating, NaN -> 0 (Rust `as`)
    if (f != f) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return -2147483647 - 1;
    return (int)f;
}

constexpr int kTile = 16, kApron = 4, kSpan = kTile + 2 * kApron;

// Film::report_samples as a gather (film.rs:78-114): pixel (x, y) visits the
// samples of pixels y-4..y+4, x-4..x+4 in row-major order -- the order the
// serial splat reaches it -- and applies each sample's own footprint test.
// A sample's footprint (the pixel range its splat covers, at most 5 x 5) and
// the Mitchell factors of its columns and rows are computed once, when the
// tile stages it in LDS, with the same arithmetic the splat uses; each pixel
// then takes its weight as the product of its column's and row's factors.
// The visit is branch-free: a sample that does not cover the pixel (or lies
// outside the film) adds +0 through a select, which leaves the sum's bits
// unchanged (the sums start at +0 and never become -0), so every wave runs
// the 81 visits at full width instead of diverging on each.
// Light-traced splats are summed as 64-bit fixed point (units of 2^-32) so
// that the per-pixel sums do not depend on the order in which the atomics
// land: a batch's image is the same on every run.  Range: |sum| < 2^31 per
// pixel and batch.
constexpr double kSplatScale = 4294967296.0;

__global__ __launch_bounds__(kTile* kTile) void k_film_gather(RenderArgs a, float4* film, uint32_t iters,
                                                              const long long* splat, const float* ew) {
    constexpr int kSpanN = kSpan * kSpan;
    // per staged sample: radiance; its footprint as four bytes, first and
    // last column and first and last row relative to the tile's corner less
    // kApron + kFoot (empty outside the film); and the footprint's Mitchell
    // factors by column and by row (stride 5: odd, so lanes reading
    // consecutive samples hit distinct banks)
    __shared__ float s_r[kSpanN], s_g[kSpanN], s_b[kSpanN];
    __shared__ uint32_t s_box[kSpanN];
    __shared__ float s_wx[kSpanN * 5], s_wy[kSpanN * 5];
    constexpr int kFoot = 8;  // footprint columns reach 2 + 2 beyond the apron
    const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const int x = x0 + tx, y = y0 + ty;
    const int W = (int)a.width, H = (int)a.height;
    const bool inside = x < W && y < H;
    float4 acc = inside ? film[(size_t)y * W + x] : make_float4(0, 0, 0, 0);
    const float wf = (float)a.width, hf = (float)a.height;
    const int last_col = W - 1, last_row = H - 1;
    const uint32_t npx = a.width * a.height;
    // Each thread stages up to kPer samples per iteration; the next
    // iteration's are fetched into registers while this one is gathered.
    constexpr int kThreads = kTile * kTile, kPer = (kSpanN + kThreads - 1) / kThreads;
    float4 pL[kPer];
    float2 pst[kPer];
    auto fetch = [&](uint32_t it) {
        const uint64_t it0 = a.ring_cap ? (a.ring_base + (uint64_t)it * npx) % a.ring_cap : (uint64_t)it * npx;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int k = (int)threadIdx.x + j * kThreads;
            const int sx = x0 - kApron + k % kSpan, sy = y0 - kApron + k / kSpan;
            if (k < kSpanN && sx >= 0 && sx < W && sy >= 0 && sy < H) {
                uint64_t q = it0 + (uint64_t)sy * W + sx;
                if (a.ring_cap && q >= a.ring_cap) q -= a.ring_cap;  // q < npx + ring_cap <= 2 ring_cap
                pst[j] = a.rb.st[(size_t)q];
                pL[j] = a.rb.L[(size_t)q];
            }
        }
    };
    if (iters) fetch(0);
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int k = (int)threadIdx.x + j * kThreads;
            if (k >= kSpanN) continue;
            const int sx = x0 - kApron + k % kSpan, sy = y0 - kApron + k / kSpan;
            if (sx >= 0 && sx < W && sy >= 0 && sy < H) {
                const float2 st = pst[j];
                const float4 L = pL[j];
                const float col_cont = 0.0f + (0.5f * (st.x + 1.0f)) * (wf - 0.0f);
                const float row_cont = 0.0f + (0.5f * (st.y + 1.0f)) * (hf - 0.0f);
                const float cd = col_cont - 0.5f, rd = row_cont - 0.5f;
                const int c0 = clampi(f2i(ceilf(cd - 2.0f)), 0, last_col);
                const int r0 = clampi(f2i(ceilf(rd - 2.0f)), 0, last_row);
                const int c1 = clampi(f2i(floorf(cd + 2.0f)), 0, last_col);
                const int r1 = clampi(f2i(floorf(rd + 2.0f)), 0, last_row);
                // the footprint of a sample staged for this tile lies within
                // kFoot of the apron (its pixel is, and the footprint spans
                // +-4 around it; a NaN position gives column / row 0, which
                // covers no pixel of a tile that far and is packed as empty)
                const int bx = x0 - kApron - kFoot, by = y0 - kApron - kFoot;
                const bool fits = c0 - bx >= 0 && c1 - bx < 256 && r0 - by >= 0 && r1 - by < 256;
                s_box[k] = fits ? (uint32_t)(c0 - bx) | (uint32_t)(c1 - bx) << 8 | (uint32_t)(r0 - by) << 16 |
                                      (uint32_t)(r1 - by) << 24
                                : 0x00010001u;  // first 1, last 0: covers no pixel
                for (int i = 0; i < 5; ++i) {
                    s_wx[k * 5 + i] = mitchell1(((float)(c0 + i) - cd) / 2.0f);
                    s_wy[k * 5 + i] = mitchell1(((float)(r0 + i) - rd) / 2.0f);
                }
                s_r[k] = L.x;
                s_g[k] = L.y;
                s_b[k] = L.z;
                // counted once, by the tile that holds the sample's own pixel
                const bool home = sx >= x0 && sx < x0 + kTile && sy >= y0 && sy < y0 + kTile;
                note_stat(a.stats, kStatNonfiniteRadiance, home && !finite3(mk(L.x, L.y, L.z)));
            } else {
                s_box[k] = 0x00010001u;  // covers no pixel
            }
        }
        __syncthreads();
        if (it + 1 < iters) fetch(it + 1);
        if (inside) {
            const int kb = (y - y0) * kSpan + (x - x0);
            // this pixel relative to the footprints' origin
            const int px = x - (x0 - kApron - kFoot), py = y - (y0 - kApron - kFoot);
#pragma unroll 1
            for (int dy = 0; dy < 2 * kApron + 1; ++dy) {
#pragma unroll
                for (int dx = 0; dx < 2 * kApron + 1; ++dx) {
                    const int k = kb + dy * kSpan + dx;
                    const uint32_t bx = s_box[k];
                    const int jx = px - (int)(bx & 0xffu), jy = py - (int)((bx >> 16) & 0xffu);
                    const bool cov = jx >= 0 && px <= (int)((bx >> 8) & 0xffu) && jy >= 0 && py <= (int)(bx >> 24);
                    const int ix = cov ? jx : 0, iy = cov ? jy : 0;
                    const float wt = s_wx[k * 5 + ix] * s_wy[k * 5 + iy];
                    acc.x = acc.x + (cov ? s_r[k] * wt : 0.0f);
                    acc.y = acc.y + (cov ? s_g[k] * wt : 0.0f);
                    acc.z = acc.z + (cov ? s_b[k] * wt : 0.0f);
                    acc.w = acc.w + (cov ? wt : 0.0f);
                }
            }
        }
        __syncthreads();
    }
    if (inside && splat) {
        // BDPT's light-traced radiance of the batch, weighted by the film weight
        // the batch's camera samples give the pixel in expectation (per iteration:
        // its column's expectation times its row's)
        const float e = ew[x] * ew[W + y];
        const long long* q = splat + 4 * ((size_t)y * W + x);
        const float4 sp = make_float4((float)((double)q[0] * (1.0 / kSplatScale)),
                                      (float)((double)q[1] * (1.0 / kSplatScale)),
                                      (float)((double)q[2] * (1.0 / kSplatScale)), 0.0f);
        acc.x = acc.x + e * sp.x;
        acc.y = acc.y + e * sp.y;
        acc.z = acc.z + e * sp.z;
    }
    if (inside) film[(size_t)y * W + x] = acc;
}

// Per-function material evaluation (test hook for gr_material_query).
__global__ void k_material_query(ShadeView sh, uint32_t prim, const float* in, uint32_t n, float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* q = in + 24 * (size_t)i;
    float* r = out + 20 * (size_t)i;
    const GpuMaterial& m = sh.mats[prim];
    const vec3 wi = mk(q[0], q[1], q[2]), wo = mk(q[3], q[4], q[5]);
    Frame f;
    f.n = mk(q[6], q[7], q[8]);
    f.t = mk(q[9], q[10], q[11]);
    f.b = mk(q[12], q[13], q[14]);
    f.gn = mk(q[15], q[16], q[17]);
    const bool c2l = q[21] != 0.0f;
    const vec3 fw = dev::f_world(m, wi, wo, f, c2l);
    const float pw = dev::pdf_world(m, wi, wo, f);
    const vec3 lw = dev::light_world(m, wi, f);
    const dev::MatSample s = dev::sample_world(m, wi, f, c2l, __float_as_uint(q[18]), q[19], q[20]);
    r[0] = fw.x; r[1] = fw.y; r[2] = fw.z; r[3] = pw;
    r[4] = lw.x; r[5] = lw.y; r[6] = lw.z;
    r[7] = s.emission.x; r[8] = s.emission.y; r[9] = s.emission.z;
    r[10] = s.radiance.x; r[11] = s.radiance.y; r[12] = s.radiance.z;
    r[13] = s.outgoing.x; r[14] = s.outgoing.y; r[15] = s.outgoing.z;
    r[16] = s.pdf; r[17] = (float)s.kind;
    r[18] = (float)dev::count_lobes(m, kLobeDiffuse | kLobeGlossy);
    r[19] = dev::sample_density_world(m, wi, wo, f);
}

// ---------------------------------------------------------------------------
// BDPT (BdptIntegrator, src/render/integrators.rs:102-418)
// ---------------------------------------------------------------------------
// A subpath vertex (BdptVertex, integrators.rs:107-130) in 7 float4: the
// fields the estimator reads.  pdf_forward / pdf_reverse are tracked by the
// reference but never enter its estimate (uniform strategy weights), so they
// are not stored.
struct BVert {
    float4 a;  // incoming.xyz, prim (bits; kNone = escaped)
    float4 b;  // point.xyz, connectible (bits)
    float4 c;  // shading normal.xyz, throughput.x
    float4 d;  // tangent.xyz, throughput.y
    float4 e;  // binormal.xyz, throughput.z
    float4 f;  // geometric normal.xyz, emission.x
    float4 g;  // emission.y, emission.z, MIS area densities: from the walk's predecessor, from its successor
};

struct BdptArgs {
    BVert* cam;   // [cmax][npaths]
    BVert* lig;   // [lmax][npaths]
    uint32_t* nc;        // camera vertex count | kLastDead
    uint32_t* nl;        // light vertex count | kLastDead, or kLightFailed
    uint32_t* mask;      // visibility bits per path, mwords words each
    uint2* entry;        // shadow queue: (path, strategy)
    uint32_t* list;      // the chunk's valid strategies, as s * chunk + (path - p0)
    uint32_t mis;        // 0: the reference's 1/(c+l); 1: power-heuristic MIS; 2: MIS + light tracing
    float4* slots;       // per (strategy, path): contribution * weight, needs-visibility flag
    long long* splat;    // light tracing: per pixel, the batch's splatted radiance (x, y, z, unused) in kSplatScale units
    float* ew;           // light tracing: expected film weight per column (width) then per row (height)
    uint32_t cmax, lmax, mwords;
    uint32_t lt_base;    // strategy index of light tracing's l = 1: cmax (lmax + 1)
};

constexpr uint32_t kBdptMaxDepth = 16;    // BDPT_MAX_DEPTH, integrators.rs:103
// Per-path subpath words nc (camera) and nl (light): a vertex count and flags.
constexpr uint32_t kVertCount = 0xffffu;        // the vertex count
constexpr uint32_t kLightFailed = 0x80000000u;  // nl: the light sample failed (no light subpath)
constexpr uint32_t kLastDead = 0x40000000u;     // the subpath's last vertex is not connectible

__device__ __forceinline__ void store_vert(BVert* v, vec3 wi, uint32_t prim, vec3 pt, bool conn, const Frame& f,
                                           vec3 T, vec3 le, float a_fwd = 0.0f) {
    v->a = make_float4(wi.x, wi.y, wi.z, __uint_as_float(prim));
    v->b = make_float4(pt.x, pt.y, pt.z, __uint_as_float(conn ? 1u : 0u));
    v->c = make_float4(f.n.x, f.n.y, f.n.z, T.x);
    v->d = make_float4(f.t.x, f.t.y, f.t.z, T.y);
    v->e = make_float4(f.b.x, f.b.y, f.b.z, T.z);
    v->f = make_float4(f.gn.x, f.gn.y, f.gn.z, le.x);
    v->g = make_float4(le.y, le.z, a_fwd, 0.0f);
}

struct VertView {
    vec3 wi, pt, T, le;
    Frame f;
    uint32_t prim;
    bool conn;
};

__device__ __forceinline__ VertView load_vert(const BVert* v) {
    VertView o;
    const float4 a = v->a, b = v->b, c = v->c, d = v->d, e = v->e, f = v->f, g = v->g;
    o.wi = mk(a.x, a.y, a.z);
    o.prim = __float_as_uint(a.w);
    o.pt = mk(b.x, b.y, b.z);
    o.conn = __float_as_uint(b.w) != 0u;
    o.f.n = mk(c.x, c.y, c.z);
    o.f.t = mk(d.x, d.y, d.z);
    o.f.b = mk(e.x, e.y, e.z);
    o.f.gn = mk(f.x, f.y, f.z);
    o.T = mk(c.w, d.w, e.w);
    o.le = mk(f.w, g.x, g.y);
    return o;
}

// correct_shading_normal, integrators.rs:232-251.
__device__ __forceinline__ float correct_shading_normal(vec3 wi, vec3 wo, const Frame& f, bool c2l) {
    if (c2l) return 1.0f;
    const float num = fabsf(dot(wi, f.n)) * fabsf(dot(wo, f.gn));
    const float den = fabsf(dot(wi, f.gn)) * fabsf(dot(wo, f.n));
    return den == 0.0f ? 0.0f : num / den;
}

// convert_density, integrators.rs:253-266.
__device__ __forceinline__ float convert_density(vec3 prev, vec3 next, const Frame& nf) {
    const float dist = length(next - prev);
    if (dist == 0.0f) return 0.0f;
    const vec3 dir = (next - prev) / dist;
    return fabsf(dot(nf.n, dir)) / (dist * dist);
}

// A solid-angle density at `from` toward `to` as an area density at `to`
// (convert_density with the receiving vertex's shading normal n).
__device__ __forceinline__ float to_area(vec3 from, vec3 to, vec3 n) {
    const float dist = length(to - from);
    if (dist == 0.0f) return 0.0f;
    const vec3 dir = (to - from) / dist;
    return fabsf(dot(n, dir)) / (dist * dist);
}

// One step of random_walk (integrators.rs:138-229) for every queued subpath.
template <bool CAMERA>
__global__ __launch_bounds__(256) void k_walk(RenderArgs a, BdptArgs b, uint32_t k, uint32_t cur) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = a.rb.counts[cur];
    bool alive = false;
    uint32_t p = 0;
    float next_pdf = 0.0f;
    vec3 new_o = mk(0.0f, 0.0f, 0.0f), new_d = new_o;
    if (i < n) {
        p = a.rb.queue[cur][i];
        const uint4 hv = a.rb.hits[i];
        const gr_hit h = {hv.x, hv.y, __uint_as_float(hv.z), hv.w};
        const vec3 o = ld3(a.rb.rays[cur][2 * (size_t)i]), d = ld3(a.rb.rays[cur][2 * (size_t)i + 1]);
        bool vconn = false;  // this vertex connectible (an escaped ray's is not)
        const float4 T4 = a.rb.T[p];
        vec3 T = ld3(T4);
        const float prev_pdf = T4.w;  // solid-angle density of d at the previous vertex
        BVert* v = (CAMERA ? b.cam : b.lig) + (size_t)k * a.npaths + p;
        if (h.prim != kNone) {
            const vec3 prev_T = T;
            const vec3 hp = o + d * h.t;
            const vec3 wi = -d;
            const Frame f = hit_frame(a.sv, a.sh, h, o, d);
            const GpuMaterial& m = a.sh.mats[h.prim];
            const uint4 r0 = draws(a.seed, a.path_base + p, CAMERA ? kStreamCamera : kStreamLight, k, 0);
            const dev::MatSample smp = dev::sample_world(m, wi, f, CAMERA, r0.x, u01(r0.y), u01(r0.z));
            note_stat(a.stats, kStatNonfiniteBsdf, bad_sample(smp));
            const bool conn = dev::count_lobes(m, kLobeDiffuse | kLobeGlossy) != 0;
            vconn = conn;
            float a_fwd = 0.0f;
            if (b.mis && k > 0) {
                // area densities for the MIS weights: this vertex from its predecessor,
                // the predecessor from this vertex (given the direction sampled here)
                BVert* pv = v - a.npaths;
                const float4 pb = pv->b, pc = pv->c;
                const vec3 pp = mk(pb.x, pb.y, pb.z), pn = mk(pc.x, pc.y, pc.z);
                a_fwd = prev_pdf * to_area(pp, hp, f.n);
                pv->g.w = dev::pdf_world(m, smp.outgoing, wi, f) * to_area(hp, pp, pn);
            } else if (b.mis) {
                a_fwd = prev_pdf * to_area(o, hp, f.n);  // the camera's first vertex, from the camera
            }
            T = mul(T, smp.radiance * (fabsf(dot(f.n, smp.outgoing)) / smp.pdf));
            T = T * correct_shading_normal(wi, smp.outgoing, f, CAMERA);
            new_o = hp + smp.outgoing * kRayPush;
            new_d = smp.outgoing;
            store_vert(v, wi, h.prim, hp, conn, f, prev_T, smp.emission, a_fwd);
            next_pdf = smp.pdf;
            if (k + 1 >= 4 || dev::nearly_zero3(T)) {
                const float live = 0.25f + dev::clamp_unit(dev::luminance(T)) * (0.75f - 0.25f);
                if (u01(r0.w) < live) T = T / live;
                else T = mk(0.0f, 0.0f, 0.0f);
            }
            alive = !dev::is_zero(T) && k + 1 < (CAMERA ? b.cmax : b.lmax);
        } else {
            // the escaped ray: a vertex that never connects (integrators.rs:210-226)
            Frame z;
            z.n = z.t = z.b = z.gn = mk(0.0f, 0.0f, 0.0f);
            T = mk(0.0f, 0.0f, 0.0f);
            store_vert(v, -d, kNone, mk(0.0f, 0.0f, 0.0f), false, z, T, T);
        }
        a.rb.T[p] = make_float4(T.x, T.y, T.z, next_pdf);
        // the last walk step to reach p writes the final word: its count and
        // whether its last vertex connects
        const uint32_t word = (k + 1) | (vconn ? 0u : kLastDead);
        if (CAMERA) b.nc[p] = word;
        else b.nl[p] = word;
    }
    const uint32_t nxt = cur ^ 1u;
    const uint32_t slot = block_append(alive, &a.rb.counts[nxt]);
    if (alive) {
        a.rb.queue[nxt][slot] = p;
        a.rb.rays[nxt][2 * (size_t)slot] = make_float4(new_o.x, new_o.y, new_o.z, 0.0f);
        a.rb.rays[nxt][2 * (size_t)slot + 1] = make_float4(new_d.x, new_d.y, new_d.z, 0.0f);
    }
}

// Light sample and the light vertex (integrators.rs:373-395): Bvh::sample_light
// with Prim::sample_ray_world's cosine direction (src/geom/prim.rs:37-51).
__global__ __launch_bounds__(256) void k_light_start(RenderArgs a, BdptArgs b) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    vec3 org = mk(0.0f, 0.0f, 0.0f), dir = org;
    if (p < a.npaths) {
        const uint64_t path = a.path_base + p;
        const LightPoint lp = sample_light_point(a.sh, draws(a.seed, path, kStreamLight, 0, 1));
        const uint4 r2 = draws(a.seed, path, kStreamLight, 0, 2);
        const vec3 dl = dev::cosine_hemisphere(u01(r2.x), u01(r2.y), false);
        const float dir_pdf = fabsf(dl.z) * kInvPi;
        vec3 t, bn;
        dev::coord_system(lp.f.gn, &t, &bn);
        const vec3 gn = lp.f.gn;
        dir = mk(t.x * dl.x + bn.x * dl.y + gn.x * dl.z, t.y * dl.x + bn.y * dl.y + gn.y * dl.z,
                 t.z * dl.x + bn.z * dl.y + gn.z * dl.z);
        if (lp.point_pdf == 0.0f || dir_pdf == 0.0f) {
            b.nl[p] = kLightFailed;
        } else {
            const vec3 init = dev::light_world(a.sh.mats[lp.prim], dir, lp.f) *
                              (fabsf(dot(gn, dir)) / (lp.point_pdf * dir_pdf));
            const vec3 one = mk(1.0f, 1.0f, 1.0f);
            store_vert(b.lig + p, mk(0.0f, 0.0f, 0.0f), lp.prim, lp.p, true, lp.f, one / lp.point_pdf,
                       mk(0.0f, 0.0f, 0.0f), lp.point_pdf);
            b.nl[p] = 1;
            a.rb.T[p] = make_float4(init.x, init.y, init.z, dir_pdf);
            org = lp.p + dir * kRayPush;
            alive = !dev::is_zero(init) && 1 < b.lmax;
        }
        for (uint32_t w = 0; w < b.mwords; ++w) b.mask[(size_t)p * b.mwords + w] = 0u;
    }
    const uint32_t slot = block_append(alive, &a.rb.counts[0]);
    if (alive) {
        a.rb.queue[0][slot] = p;
        a.rb.rays[0][2 * (size_t)slot] = make_float4(org.x, org.y, org.z, 0.0f);
        a.rb.rays[0][2 * (size_t)slot + 1] = make_float4(dir.x, dir.y, dir.z, 0.0f);
    }
}

struct Connection {
    vec3 contrib;
    bool needs_vis;
    vec3 from, to;
};

// BdptIntegrator::connect without the visibility test (integrators.rs:269-343).
__device__ Connection connect(const RenderArgs& a, const BdptArgs& b, uint32_t p, uint32_t c, uint32_t l) {
    Connection r;
    r.contrib = mk(0.0f, 0.0f, 0.0f);
    r.needs_vis = false;
    const VertView cv = load_vert(b.cam + (size_t)(c - 1) * a.npaths + p);
    if (l == 0) {
        if (cv.prim != kNone) r.contrib = mul(cv.T, cv.le);
        return r;
    }
    const VertView lv = load_vert(b.lig + (size_t)(l - 1) * a.npaths + p);
    if (cv.prim == kNone || lv.prim == kNone) return r;
    if (!cv.conn || !lv.conn) return r;
    const vec3 c2l = normalized(lv.pt - cv.pt);
    const vec3 l2c = -c2l;
    const vec3 cr = dev::f_world(a.sh.mats[cv.prim], cv.wi, c2l, cv.f, true);
    const vec3 ce = l == 1 ? dev::light_world(a.sh.mats[lv.prim], l2c, lv.f)
                           : dev::f_world(a.sh.mats[lv.prim], lv.wi, l2c, lv.f, false);
    const float g = fabsf(dot(cv.f.n, c2l));
    const float cd = convert_density(cv.pt, lv.pt, lv.f);
    r.contrib = mul(mul(mul(cv.T, cr), ce), lv.T) * (g * cd);
    r.needs_vis = !dev::nearly_zero3(r.contrib);
    r.from = cv.pt;
    r.to = lv.pt;
    return r;
}

// BSDF densities (solid angle) at the two joined vertices that the MIS
// weight of a strategy with c >= 1 and l >= 1 needs, in both directions.
struct JoinDensities {
    float c_fwd, c_rev;  // at x_c: toward y_l given x_c's incoming direction, and back
    float l_fwd, l_rev;  // at y_l (l >= 2): toward x_c given y_l's incoming direction, and back
};

// connect() with its two BSDF evaluations fused with the densities the MIS
// weight needs at the same vertices (same values as the separate calls).
__device__ Connection connect_mis(const RenderArgs& a, const BdptArgs& b, uint32_t p, uint32_t c, uint32_t l,
                                  JoinDensities* jd) {
    if (l == 0) return connect(a, b, p, c, l);
    Connection r;
    r.contrib = mk(0.0f, 0.0f, 0.0f);
    r.needs_vis = false;
    const VertView cv = load_vert(b.cam + (size_t)(c - 1) * a.npaths + p);
    const VertView lv = load_vert(b.lig + (size_t)(l - 1) * a.npaths + p);
    if (cv.prim == kNone || lv.prim == kNone) return r;
    if (!cv.conn || !lv.conn) return r;
    const vec3 c2l = normalized(lv.pt - cv.pt);
    const vec3 l2c = -c2l;
    const dev::FPdf fc = dev::f_pdf_world(a.sh.mats[cv.prim], cv.wi, c2l, cv.f, true);
    jd->c_fwd = fc.pdf_fwd;
    jd->c_rev = fc.pdf_rev;
    vec3 ce;
    float l_fwd = 0.0f, l_rev = 0.0f;
    if (l == 1) {
        ce = dev::light_world(a.sh.mats[lv.prim], l2c, lv.f);
    } else {
        const dev::FPdf fl = dev::f_pdf_world(a.sh.mats[lv.prim], lv.wi, l2c, lv.f, false);
        ce = fl.f;
        l_fwd = fl.pdf_fwd;
        l_rev = fl.pdf_rev;
    }
    jd->l_fwd = l_fwd;
    jd->l_rev = l_rev;
    const float g = fabsf(dot(cv.f.n, c2l));
    const float cd = convert_density(cv.pt, lv.pt, lv.f);
    r.contrib = mul(mul(mul(cv.T, fc.f), ce), lv.T) * (g * cd);
    r.needs_vis = !dev::nearly_zero3(r.contrib);
    r.from = cv.pt;
    r.to = lv.pt;
    return r;
}

__device__ __forceinline__ float remap0(float p) { return p != 0.0f ? p : 1.0f; }

// Density of a light's cosine emission (Prim::sample_ray_world,
// src/geom/prim.rs:37-51) toward dir from the light point of frame f.
__device__ __forceinline__ float emit_dir_pdf(const Frame& f, vec3 dir) {
    const float c = dot(f.gn, dir);
    return c > 0.0f ? c * kInvPi : 0.0f;
}

__device__ __forceinline__ bool vert_conn(const BVert* v) { return __float_as_uint(v->b.w) != 0u; }

// The pinhole camera as the light-tracing strategy sees it: its position
// (the camera transform of the origin), the rows of the transform's 3x3 part
// (camera-space axes in world space), the window and the film size.
struct LtCamera {
    vec3 e, ax, ay, az;
    float wx, wy;
    uint32_t width, height;
};

__device__ __forceinline__ LtCamera lt_camera(const RenderArgs& a) {
    const float* m = a.cam;
    LtCamera c;
    c.e = mk(m[12] / m[15], m[13] / m[15], m[14] / m[15]);
    c.ax = mk(m[0], m[1], m[2]);
    c.ay = mk(m[4], m[5], m[6]);
    c.az = mk(m[8], m[9], m[10]);
    c.wx = a.wx;
    c.wy = a.wy;
    c.width = a.width;
    c.height = a.height;
    return c;
}

// Projection of world point y seen from the camera: its film position
// (col_cont - 0.5, row_cont - 0.5, as the film computes it for a sample), the
// unit direction from y to the camera and cos theta.  False when y is behind
// the camera or its splat (the pixels within the filter width) misses the film.
__device__ __forceinline__ bool lt_project(const LtCamera& c, vec3 y, float& cd, float& rd, vec3& to_cam,
                                           float& cos_t) {
    const vec3 d = y - c.e;
    const float len = length(d);
    if (!(len > 0.0f)) return false;
    const vec3 u = d / len;
    const float dx = dot(u, c.ax), dy = dot(u, c.ay), dz = dot(u, c.az);
    if (!(dz < 0.0f)) return false;
    const float s = (dx / -dz) / c.wx, t = (dy / -dz) / c.wy;
    const float wf = (float)c.width, hf = (float)c.height;
    const float col_cont = 0.0f + (0.5f * (s + 1.0f)) * (wf - 0.0f);
    const float row_cont = 0.0f + (0.5f * (t + 1.0f)) * (hf - 0.0f);
    cd = col_cont - 0.5f;
    rd = row_cont - 0.5f;
    if (!(floorf(cd + 2.0f) >= 0.0f && ceilf(cd - 2.0f) <= wf - 1.0f && floorf(rd + 2.0f) >= 0.0f &&
          ceilf(rd - 2.0f) <= hf - 1.0f))
        return false;
    to_cam = -u;
    cos_t = -dz / sqrtf(dx * dx + dy * dy + dz * dz);
    return true;
}

// Power-heuristic MIS weight of strategy (c, l) of path p among the
// strategies that can make the same path z_1..z_k (k = c + l; z_i = camera
// vertex i for i <= c, z_{c+m} = light vertex l - m + 1): j camera and k - j
// light vertices, j = 1..k, within the depth caps, joining two connectible
// vertices (j < k) or ending on the emitter (j = k), with z_k on a light.
// With p_j = prod_{i<=j} pF(z_i) prod_{i>j} pL(z_i) (area densities: pF along
// the camera walk, pL along the light walk), the weight is p_c^2 / sum_j p_j^2,
// summed through the ratios p_{j-1}/p_j = pL(z_j)/pF(z_j); zero densities
// (delta lobes) count as 1.  The walks store each vertex's own densities
// (the camera's first vertex: the camera's ray density); the ones that change
// with the connection -- at the two joined vertices and their neighbours --
// are recomputed here.  With light tracing (b.mis == 2) j = 0 is a strategy
// too (all k vertices from the light walk, z_1 joined to the camera, which
// counts as connectible), and c = 0 is allowed: z_1 = y_l seen by the camera.
__device__ float mis_weight(const RenderArgs& a, const BdptArgs& b, uint32_t p, uint32_t c, uint32_t l,
                            bool have_jd = false, JoinDensities jd = JoinDensities{}) {
    const size_t np = a.npaths;
    const BVert* cam = b.cam + p;  // camera vertex i at cam[(i - 1) * np]
    const BVert* lig = b.lig + p;  // light vertex m at lig[(m - 1) * np]
    const uint32_t k = c + l;
    float pl_c = 0.0f, pl_c1 = 0.0f, pf_c1 = 0.0f, pf_c2 = 0.0f;
    bool conn_c = true, conn_c1 = false;  // conn of z_c (the camera when c = 0) and z_{c+1}
    if (c == 0) {
        const VertView yl = load_vert(lig + (size_t)(l - 1) * np);
        const LtCamera lc = lt_camera(a);
        float cd, rd, cos_t;
        vec3 d_yc;
        if (!lt_project(lc, yl.pt, cd, rd, d_yc, cos_t)) return 0.0f;
        pf_c1 = camera_dir_pdf(lc.wx, lc.wy, cos_t) * to_area(lc.e, yl.pt, yl.f.n);
        if (l >= 2) {
            const GpuMaterial& my = a.sh.mats[yl.prim];
            const float4 qb = lig[(size_t)(l - 2) * np].b, qc = lig[(size_t)(l - 2) * np].c;
            pf_c2 = dev::pdf_world(my, d_yc, yl.wi, yl.f) * to_area(yl.pt, mk(qb.x, qb.y, qb.z), mk(qc.x, qc.y, qc.z));
        }
        conn_c1 = yl.conn;
    } else {
        const VertView xc = load_vert(cam + (size_t)(c - 1) * np);
        vec3 xp = mk(0.0f, 0.0f, 0.0f), xpn = xp;  // camera vertex c - 1
        if (c >= 2) {
            const float4 pb = cam[(size_t)(c - 2) * np].b, pc = cam[(size_t)(c - 2) * np].c;
            xp = mk(pb.x, pb.y, pb.z);
            xpn = mk(pc.x, pc.y, pc.z);
        }
        if (l == 0) {
            if (xc.prim == kNone || !a.sh.mats[xc.prim].has_light) return 0.0f;
            pl_c = light_area_pdf(a.sh, xc.prim);
            conn_c = true;  // as a light point it can be connected to
            if (c >= 2) pl_c1 = emit_dir_pdf(xc.f, normalized(xp - xc.pt)) * to_area(xc.pt, xp, xpn);
        } else {
            const VertView yl = load_vert(lig + (size_t)(l - 1) * np);
            const GpuMaterial& mc = a.sh.mats[xc.prim];
            const vec3 d_cy = normalized(yl.pt - xc.pt), d_yc = -d_cy;
            pf_c1 = (have_jd ? jd.c_fwd : dev::pdf_world(mc, xc.wi, d_cy, xc.f)) * to_area(xc.pt, yl.pt, yl.f.n);
            if (l == 1) {
                pl_c = emit_dir_pdf(yl.f, d_yc) * to_area(yl.pt, xc.pt, xc.f.n);
            } else {
                const GpuMaterial& my = a.sh.mats[yl.prim];
                pl_c = (have_jd ? jd.l_fwd : dev::pdf_world(my, yl.wi, d_yc, yl.f)) * to_area(yl.pt, xc.pt, xc.f.n);
                const float4 qb = lig[(size_t)(l - 2) * np].b, qc = lig[(size_t)(l - 2) * np].c;
                pf_c2 = (have_jd ? jd.l_rev : dev::pdf_world(my, d_yc, yl.wi, yl.f)) *
                        to_area(yl.pt, mk(qb.x, qb.y, qb.z), mk(qc.x, qc.y, qc.z));
            }
            if (c >= 2) pl_c1 = (have_jd ? jd.c_rev : dev::pdf_world(mc, d_cy, xc.wi, xc.f)) * to_area(xc.pt, xp, xpn);
            conn_c = xc.conn;
            conn_c1 = yl.conn;
        }
    }
    auto valid = [&](int j, bool cj, bool cj1) {
        return (uint32_t)j <= b.cmax && k - (uint32_t)j <= b.lmax && ((uint32_t)j == k || (cj && cj1));
    };
    float sum = valid((int)c, conn_c, conn_c1) ? 1.0f : 0.0f;
    float r = 1.0f;
    bool hi = conn_c;  // conn of z_{j+1}
    const int jmin = b.mis == 2 ? 0 : 1;
    for (int j = (int)c - 1; j >= jmin; --j) {
        const BVert* vi = cam + (size_t)j * np;  // z_{j+1}
        const float pf = vi->g.z;
        const float pl = j + 1 == (int)c ? pl_c : (j + 2 == (int)c ? pl_c1 : vi->g.w);
        r = r * (remap0(pl) / remap0(pf));
        const bool cj = j == 0 || vert_conn(cam + (size_t)(j - 1) * np);
        if (valid(j, cj, hi)) sum += r * r;
        hi = cj;
    }
    r = 1.0f;
    for (uint32_t m = 1; m <= l; ++m) {
        const BVert* y = lig + (size_t)(l - m) * np;  // z_{c+m}
        const float4 g = y->g;
        const float pf = m == 1 ? pf_c1 : (m == 2 ? pf_c2 : g.w);
        r = r * (remap0(pf) / remap0(g.z));
        const bool cj = vert_conn(y);
        const bool cj1 = m < l && vert_conn(lig + (size_t)(l - m - 1) * np);
        if (valid((int)(c + m), cj, cj1)) sum += r * r;
    }
    return sum > 0.0f ? 1.0f / sum : 0.0f;
}

// The light-tracing strategy (c = 0) for light vertex l of path p: the
// radiance the light subpath carries to y and y scatters toward the camera,
// T f |n.w| / d^2, times the camera's density of that direction (the
// importance of a pinhole whose film is sampled uniformly); the visibility
// segment runs from the camera to y.
__device__ Connection light_trace(const RenderArgs& a, const BdptArgs& b, uint32_t p, uint32_t l) {
    Connection r;
    r.contrib = mk(0.0f, 0.0f, 0.0f);
    r.needs_vis = false;
    const VertView y = load_vert(b.lig + (size_t)(l - 1) * a.npaths + p);
    if (y.prim == kNone || !y.conn) return r;
    const LtCamera lc = lt_camera(a);
    float cd, rd, cos_t;
    vec3 to_cam;
    if (!lt_project(lc, y.pt, cd, rd, to_cam, cos_t)) return r;
    const GpuMaterial& m = a.sh.mats[y.prim];
    const vec3 f = l == 1 ? dev::light_world(m, to_cam, y.f) : dev::f_world(m, y.wi, to_cam, y.f, false);
    const float g = convert_density(lc.e, y.pt, y.f);
    r.contrib = mul(y.T, f) * (g * camera_dir_pdf(lc.wx, lc.wy, cos_t));
    if (dev::nearly_zero3(r.contrib)) {
        r.contrib = mk(0.0f, 0.0f, 0.0f);
        return r;
    }
    r.needs_vis = true;
    r.from = lc.e;
    r.to = y.pt;
    return r;
}

// Fixed-point value of a splatted radiance component (0 for a non-finite one).
__device__ __forceinline__ unsigned long long splat_fixed(float v) {
    if (!isfinite(v)) return 0ull;
    const double d = rint((double)v * kSplatScale);
    const double lim = 9.2e18;
    return (unsigned long long)(long long)(d > lim ? lim : (d < -lim ? -lim : d));
}

// Splats the visible light-tracing contributions of paths [p0, p1) onto the
// batch's splat buffer with the film's Mitchell weights over the pixels
// within the filter width (the footprint a camera sample at the same film
// position has), as fixed-point integer atomics (order-independent sums).
__global__ __launch_bounds__(256) void k_lt_splat(RenderArgs a, BdptArgs b, uint32_t p0, uint32_t p1) {
    const uint32_t p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= p1) return;
    const uint32_t wl = b.nl[p];
    if (wl & kLightFailed) return;
    const uint32_t nl = wl & kVertCount;
    const uint32_t chunk = p1 - p0, q = p - p0;
    const LtCamera lc = lt_camera(a);
    const int W = (int)a.width, H = (int)a.height;
    for (uint32_t l = 1; l <= nl; ++l) {
        const uint32_t s = b.lt_base + l - 1;
        const float4 v = b.slots[(size_t)s * chunk + q];
        if (v.w == 0.0f || !(b.mask[(size_t)p * b.mwords + s / 32] & (1u << (s % 32)))) continue;
        // a non-finite contribution is dropped; one beyond the fixed-point range
        // (|v| wt * 2^32 > 9.2e18 with wt <= 1) is clamped
        if (!finite3(mk(v.x, v.y, v.z))) {
            note_stat(a.stats, kStatBadSplats, true);
            continue;
        }
        note_stat(a.stats, kStatBadSplats, fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z))) > 2.0e9f);
        const float4 pb = b.lig[(size_t)(l - 1) * a.npaths + p].b;
        float cd, rd, cos_t;
        vec3 to_cam;
        if (!lt_project(lc, mk(pb.x, pb.y, pb.z), cd, rd, to_cam, cos_t)) continue;
        const int x0 = clampi(f2i(ceilf(cd - 2.0f)), 0, W - 1), x1 = clampi(f2i(floorf(cd + 2.0f)), 0, W - 1);
        const int y0 = clampi(f2i(ceilf(rd - 2.0f)), 0, H - 1), y1 = clampi(f2i(floorf(rd + 2.0f)), 0, H - 1);
        for (int y = y0; y <= y1; ++y) {
            const float wy = mitchell1(((float)y - rd) / 2.0f);
            for (int x = x0; x <= x1; ++x) {
                const float wt = mitchell1(((float)x - cd) / 2.0f) * wy;
                unsigned long long* px = reinterpret_cast<unsigned long long*>(b.splat + 4 * ((size_t)y * W + x));
                atomicAdd(px + 0, splat_fixed(v.x * wt));
                atomicAdd(px + 1, splat_fixed(v.y * wt));
                atomicAdd(px + 2, splat_fixed(v.z * wt));
            }
        }
    }
}

// Antiderivative of the 1-D Mitchell kernel k(t) = mitchell_filter1(t / 2)
// (B = C = 1/3, support |t| <= 2, integral 1), odd, in double.
__device__ double mitchell_integral(double t) {
    const double a = fabs(t);
    double r;
    if (a <= 1.0) {
        r = (1.75 * a * a * a * a - 4.0 * a * a * a + (16.0 / 3.0) * a) / 6.0;
    } else {
        const double u = a < 2.0 ? a : 2.0;
        const double fu = -(7.0 / 12.0) * u * u * u * u + 4.0 * u * u * u - 10.0 * u * u + (32.0 / 3.0) * u;
        const double f1 = -(7.0 / 12.0) * 1.0 * 1.0 * 1.0 * 1.0 + 4.0 * 1.0 * 1.0 * 1.0 - 10.0 * 1.0 * 1.0 +
                          (32.0 / 3.0) * 1.0;
        r = (1.75 - 4.0 + 16.0 / 3.0) / 6.0 + (fu - f1) / 6.0;
    }
    return t < 0.0 ? -r : r;
}

// Expected film weight one iteration's samples give each pixel column
// (threads [0, width)) and row (threads [width, width + height)): the sample
// of pixel q lies uniformly within q +- 2, so E = (1/4) sum_q
// int_{-2}^{2} k(x - q - j) dj -- 1 inside the film, less within two pixels of
// its edges.  A pixel's expectation is its column's times its row's.
__global__ void k_expected_weights(uint32_t width, uint32_t height, float* ew) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= width + height) return;
    const int64_t n = i < width ? width : height;
    const int64_t x = i < width ? i : i - width;
    double e = 0.0;
    for (int64_t q = x - 4; q <= x + 4; ++q) {
        if (q < 0 || q >= n) continue;
        e += mitchell_integral((double)(x - q) + 2.0) - mitchell_integral((double)(x - q) - 2.0);
    }
    ew[i] = (float)(0.25 * e);
}

__global__ __launch_bounds__(256) void k_vis_scatter(RenderArgs a, BdptArgs b) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.rb.counts[2] || a.rb.vis[j] == 0) return;
    const uint2 e = b.entry[j];
    atomicOr(&b.mask[(size_t)e.x * b.mwords + e.y / 32], 1u << (e.y % 32));
}

// Sums the strategies of paths [p0, p1) in the reference's order
// (integrators.rs:401-412): light += connect(c, l) * weight(c, l), with the
// strategies whose visibility test failed contributing zero.
__global__ __launch_bounds__(256) void k_connect_sum(RenderArgs a, BdptArgs b, uint32_t p0, uint32_t p1) {
    const uint32_t p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= p1) return;
    const uint32_t chunk = p1 - p0, q = p - p0;
    const uint32_t wl = b.nl[p];
    uint32_t nc = b.nc[p] & kVertCount;
    const uint32_t nl = wl & kVertCount;
    vec3 L = mk(0.0f, 0.0f, 0.0f);
    if (wl & kLightFailed) nc = 0;
    for (uint32_t c = 1; c < nc + 1; ++c) {
        for (uint32_t l = 0; l < nl + 1; ++l) {
            if (c + l > kBdptMaxDepth) continue;
            const uint32_t s = (c - 1) * (b.lmax + 1) + l;
            const float4 v = b.slots[(size_t)s * chunk + q];
            vec3 cw = mk(v.x, v.y, v.z);
            if (v.w != 0.0f && !(b.mask[(size_t)p * b.mwords + s / 32] & (1u << (s % 32))))
                cw = mk(0.0f, 0.0f, 0.0f);  // an occluded connection contributes zero * weight
            L = L + cw;
        }
    }
    a.rb.L[p] = make_float4(L.x, L.y, L.z, 0.0f);
}

inline dim3 blocks_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// Grows the film's wavefront buffers to hold `paths` paths.
hipError_t reserve(RenderBuffers& rb, size_t paths) {
    if (rb.capacity >= paths) return hipSuccess;
    (void)hipFree(rb.blob);
    rb.blob = nullptr;
    rb.capacity = 0;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off = 0;
    size_t o_rays0 = off; off = align(off + paths * 32);
    size_t o_rays1 = off; off = align(off + paths * 32);
    size_t o_q0 = off; off = align(off + paths * 4);
    size_t o_q1 = off; off = align(off + paths * 4);
    size_t o_hits = off; off = align(off + paths * 16);
    size_t o_T = off; off = align(off + paths * 16);
    size_t o_L = off; off = align(off + paths * 16);
    size_t o_aux = off; off = align(off + paths * 16);
    size_t o_st = off; off = align(off + paths * 8);
    size_t o_seg = off; off = align(off + paths * 32);
    size_t o_con = off; off = align(off + paths * 16);
    size_t o_sp = off; off = align(off + paths * 4);
    size_t o_vis = off; off = align(off + paths * 4);
    size_t o_cnt = off; off = align(off + 64);
    char* b = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&b), off);
    if (e != hipSuccess) return e;
    rb.blob = b;
    rb.capacity = paths;
    rb.rays[0] = reinterpret_cast<float4*>(b + o_rays0);
    rb.rays[1] = reinterpret_cast<float4*>(b + o_rays1);
    rb.queue[0] = reinterpret_cast<uint32_t*>(b + o_q0);
    rb.queue[1] = reinterpret_cast<uint32_t*>(b + o_q1);
    rb.hits = reinterpret_cast<uint4*>(b + o_hits);
    rb.T = reinterpret_cast<float4*>(b + o_T);
    rb.L = reinterpret_cast<float4*>(b + o_L);
    rb.aux = reinterpret_cast<float4*>(b + o_aux);
    rb.st = reinterpret_cast<float2*>(b + o_st);
    rb.seg = reinterpret_cast<float4*>(b + o_seg);
    rb.contrib = reinterpret_cast<float4*>(b + o_con);
    rb.seg_path = reinterpret_cast<uint32_t*>(b + o_sp);
    rb.vis = reinterpret_cast<uint32_t*>(b + o_vis);
    rb.counts = reinterpret_cast<uint32_t*>(b + o_cnt);
    return hipSuccess;
}

// BDPT buffers: vertices for cmax camera and lmax light vertices per path,
// counts, visibility masks, and a shadow queue of `shadow` entries.
hipError_t reserve_bdpt(RenderBuffers& rb, size_t npaths, uint32_t cmax, uint32_t lmax, uint32_t nstrat,
                        size_t shadow, size_t nslots, size_t npx, size_t nfilm_lines, BdptArgs* b, float4** seg,
                        uint32_t** vis) {
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const uint32_t mwords = (nstrat + 31) / 32;
    size_t off = 0;
    const size_t o_cam = off; off = align(off + (size_t)cmax * npaths * sizeof(BVert));
    const size_t o_lig = off; off = align(off + (size_t)lmax * npaths * sizeof(BVert));
    const size_t o_nc = off; off = align(off + npaths * 4);
    const size_t o_nl = off; off = align(off + npaths * 4);
    const size_t o_mask = off; off = align(off + npaths * mwords * 4);
    const size_t o_entry = off; off = align(off + shadow * sizeof(uint2));
    const size_t o_seg = off; off = align(off + shadow * 32);
    const size_t o_vis = off; off = align(off + shadow * 4);
    const size_t o_slots = off; off = align(off + nslots * sizeof(float4));
    const size_t o_list = off; off = align(off + nslots * sizeof(uint32_t));
    const size_t o_splat = off; off = align(off + npx * 4 * sizeof(long long));
    const size_t o_ew = off; off = align(off + nfilm_lines * sizeof(float));
    if (rb.bdpt_capacity < off) {
        (void)hipFree(rb.bdpt_blob);
        rb.bdpt_blob = nullptr;
        rb.bdpt_capacity = 0;
        hipError_t e = hipMalloc(&rb.bdpt_blob, off);
        if (e != hipSuccess) return e;
        rb.bdpt_capacity = off;
    }
    char* base = static_cast<char*>(rb.bdpt_blob);
    b->cam = reinterpret_cast<BVert*>(base + o_cam);
    b->lig = reinterpret_cast<BVert*>(base + o_lig);
    b->nc = reinterpret_cast<uint32_t*>(base + o_nc);
    b->nl = reinterpret_cast<uint32_t*>(base + o_nl);
    b->mask = reinterpret_cast<uint32_t*>(base + o_mask);
    b->entry = reinterpret_cast<uint2*>(base + o_entry);
    b->slots = reinterpret_cast<float4*>(base + o_slots);
    b->list = reinterpret_cast<uint32_t*>(base + o_list);
    b->splat = reinterpret_cast<long long*>(base + o_splat);
    b->ew = reinterpret_cast<float*>(base + o_ew);
    b->lt_base = cmax * (lmax + 1);
    b->cmax = cmax;
    b->lmax = lmax;
    b->mwords = mwords;
    *seg = reinterpret_cast<float4*>(base + o_seg);
    *vis = reinterpret_cast<uint32_t*>(base + o_vis);
    return hipSuccess;
}

}  // namespace

namespace {
constexpr uint32_t kMaxStrategies = kBdptMaxDepth * (kBdptMaxDepth + 1);  // cmax (lmax + 1) at most

// Lists the valid (strategy, path) pairs of a 256-thread block's paths,
// strategy by strategy: the pairs of strategy s = i * ni + j (i < no, j < ni)
// sit together in thread order, so a wave of the evaluation runs one
// strategy over neighbouring paths -- uniform branches, neighbouring vertex
// records, slots written side by side.  One atomic per block reserves the
// block's run of `list` from `counter`; entry = (s0 + s) * chunk + q.  Every
// thread of the block calls it; valid(i, j) is false for paths past the chunk.
template <typename Valid>
__device__ void list_by_strategy(uint32_t no, uint32_t ni, Valid valid, uint32_t q, uint32_t chunk, uint32_t s0,
                                 uint32_t* counter, uint32_t* list) {
    __shared__ uint32_t s_wave[4][kMaxStrategies];  // per wave, then the wave's offset within its strategy
    __shared__ uint32_t s_tot[kMaxStrategies];      // per strategy, then the strategy's offset in the run
    __shared__ uint32_t s_first;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t ns = no * ni;
    for (uint32_t i = 0, s = 0; i < no; ++i)
        for (uint32_t j = 0; j < ni; ++j, ++s) {
            const uint64_t m = __ballot(valid(i, j));
            if (lane == 0) s_wave[wave][s] = (uint32_t)__popcll(m);
        }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < ns; s += blockDim.x) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < 4; ++w) {
            const uint32_t c = s_wave[w][s];
            s_wave[w][s] = run;
            run += c;
        }
        s_tot[s] = run;
    }
    __syncthreads();
    if (wave == 0) {  // exclusive scan of the strategy totals, 64 at a time
        uint32_t carry = 0;
        for (uint32_t base = 0; base < ns; base += 64) {
            const uint32_t s = base + lane;
            const uint32_t v = s < ns ? s_tot[s] : 0u;
            uint32_t incl = v;
#pragma unroll
            for (uint32_t d = 1; d < 64; d <<= 1) {
                const uint32_t u = __shfl_up(incl, d);
                if (lane >= d) incl += u;
            }
            if (s < ns) s_tot[s] = carry + incl - v;
            carry += __shfl(incl, 63);
        }
        if (lane == 0) s_first = carry ? atomicAdd(counter, carry) : 0u;
    }
    __syncthreads();
    const uint32_t first = s_first;
    for (uint32_t i = 0, s = 0; i < no; ++i)
        for (uint32_t j = 0; j < ni; ++j, ++s) {
            const bool v = valid(i, j);
            const uint64_t m = __ballot(v);
            if (v) list[first + s_tot[s] + s_wave[wave][s] + (uint32_t)__popcll(m & lanes_below())] = (s0 + s) * chunk + q;
        }
}

// The strategies (c, l) of paths [p0, p1) that connect() evaluates -- c <=
// nc, l <= nl, c + l <= BDPT_MAX_DEPTH, light sample not failed -- listed as
// s * chunk + (p - p0), s = (c - 1)(lmax + 1) + l.  Most of the cmax (lmax + 1)
// strategies of a path are invalid (subpaths end early), so evaluating only
// the listed ones keeps every lane of a wave busy.  Light tracing's
// strategies (0, l) go to a list of their own (after the connection
// strategies' region), so that every wave of an evaluation takes one branch.
__global__ __launch_bounds__(256) void k_connect_list(RenderArgs a, BdptArgs b, uint32_t p0, uint32_t p1) {
    const uint32_t chunk = p1 - p0;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nc = 0, nl = 0, nlt = 0;
    uint32_t c_dead = kNone, l_dead = kNone;  // a subpath's last vertex when it is not connectible
    if (q < chunk) {
        const uint32_t wl = b.nl[p0 + q];
        if (!(wl & kLightFailed)) {
            const uint32_t wc = b.nc[p0 + q];
            nc = wc & kVertCount;
            nl = wl & kVertCount;
            if (wc & kLastDead) c_dead = nc;
            if (wl & kLastDead) l_dead = nl;
            if (b.mis == 2) nlt = nl;  // light tracing: (0, l), l = 1..nl
        }
    }
    // A strategy joining a non-connectible vertex (an escaped ray's, one on a
    // light) contributes zero without a visibility test: its slot is written
    // here and it is not listed, so the evaluation's waves hold live
    // strategies only.  Only a subpath's last vertex is checked (that is
    // where escapes and light hits end a walk).
    for (uint32_t c = 1; c <= nc; ++c) {
        const uint32_t lm = min(nl, kBdptMaxDepth - c);
        for (uint32_t l = 1; l <= lm; ++l)
            if (c == c_dead || l == l_dead)
                b.slots[(size_t)((c - 1) * (b.lmax + 1) + l) * chunk + q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (l_dead != kNone && l_dead <= nlt)
        b.slots[(size_t)(b.lt_base + l_dead - 1) * chunk + q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    list_by_strategy(
        b.cmax, b.lmax + 1,
        [&](uint32_t i, uint32_t l) {
            return i + 1 <= nc && l <= min(nl, kBdptMaxDepth - (i + 1)) && (l == 0 || (i + 1 != c_dead && l != l_dead));
        },
        q, chunk, 0u, &a.rb.counts[4], b.list);
    if (b.mis == 2)
        list_by_strategy(
            1u, b.lmax, [&](uint32_t, uint32_t j) { return j + 1 <= nlt && j + 1 != l_dead; }, q, chunk, b.lt_base,
            &a.rb.counts[5], b.list + (size_t)b.lt_base * chunk);
}

// One lane per listed strategy (grid-stride over the list, whole blocks per
// step): evaluates connect(c, l) -- or, over light tracing's list (LT), the
// light vertex seen by the camera -- without its visibility test, stores
// contribution * weight in the strategy's slot, and queues the shadow segment
// of every strategy that needs one (wave-compacted).  One instance per list
// keeps each instance's registers to its own branch.  (A separate instance
// for the reference's weight without the MIS branch measured slower: its
// compiled form keeps a private stack object.)
template <bool LT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_connect_eval(RenderArgs a, BdptArgs b,
                                                                                            uint32_t p0, uint32_t p1) {
    const uint32_t chunk = p1 - p0;
    const uint32_t n = a.rb.counts[LT ? 5 : 4];
    const uint32_t* list = b.list + (LT ? (size_t)b.lt_base * chunk : 0);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {  // block-uniform steps
        const uint32_t j = base + threadIdx.x;
        bool emit = false;
        uint32_t p = 0, s = 0;
        Connection cn;
        if (j < n) {
            const uint32_t t = list[j];
            s = t / chunk;
            p = p0 + (t - s * chunk);
            vec3 cw;
            if (LT) {  // light tracing: light vertex l seen by the camera
                const uint32_t l = s - b.lt_base + 1;
                cn = light_trace(a, b, p, l);
                cw = dev::is_zero(cn.contrib) ? cn.contrib : cn.contrib * mis_weight(a, b, p, 0, l);
                emit = cn.needs_vis;
            } else {
                const uint32_t c = s / (b.lmax + 1) + 1, l = s % (b.lmax + 1);
                float w;
                if (!b.mis) {
                    cn = connect(a, b, p, c, l);
                    w = 1.0f / (float)(c + l);  // weight, integrators.rs:345-354
                } else {
                    JoinDensities jd{};  // set by connect_mis when l > 0
                    cn = connect_mis(a, b, p, c, l, &jd);
                    w = dev::is_zero(cn.contrib) ? 0.0f : mis_weight(a, b, p, c, l, l > 0, jd);
                }
                cw = cn.contrib * w;
                emit = l > 0 && cn.needs_vis;
            }
            b.slots[t] = make_float4(cw.x, cw.y, cw.z, emit ? 1.0f : 0.0f);
        }
        const uint32_t slot = block_append(emit, &a.rb.counts[2]);
        if (emit) {
            a.rb.seg[2 * (size_t)slot] = make_float4(cn.from.x, cn.from.y, cn.from.z, 0.0f);
            a.rb.seg[2 * (size_t)slot + 1] = make_float4(cn.to.x, cn.to.y, cn.to.z, 0.0f);
            b.entry[slot] = make_uint2(p, s);
        }
    }
}

// BdptIntegrator::integrate for every path of one batch, then the film splat.
hipError_t bdpt_batch(gr_scene* sc, RenderArgs a, gr_film* film, uint32_t cmax, uint32_t lmax, uint32_t nb,
                      uint32_t mis, float4* sample_out, hipStream_t s) {
    const size_t npaths = a.npaths;
    // strategies (c, l), c >= 1, at s = (c - 1)(lmax + 1) + l; light tracing's (0, l) after them
    const size_t nstrat = (size_t)cmax * (lmax + 1) + (mis == 2 ? lmax : 0);
    // strategies per connection chunk: at most 2^27 (2 GiB of slots, 512 MiB of
    // strategy list)
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(npaths, ((size_t)1 << 27) / nstrat));
    // segments: (c, l) with l >= 1, plus light tracing's (0, l)
    const size_t shadow = chunk * (cmax + (mis == 2 ? 1 : 0)) * lmax;
    BdptArgs b;
    float4* seg = nullptr;
    uint32_t* vis = nullptr;
    const size_t npx = (size_t)a.width * a.height;
    hipError_t e = reserve_bdpt(film->rb, npaths, cmax, lmax, (uint32_t)nstrat, shadow, chunk * nstrat, npx,
                                (size_t)a.width + a.height, &b, &seg, &vis);
    if (e == hipSuccess && sc->tuning.sort_bits) e = reserve_sort(sc->sort, std::max(npaths, shadow));
    if (e != hipSuccess) return e;
    b.mis = mis;
    if (mis == 2) {
        e = hipMemsetAsync(b.splat, 0, npx * 4 * sizeof(long long), s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_expected_weights, blocks_for(a.width + a.height), dim3(256), 0, s, a.width, a.height,
                           b.ew);
    }
    a.rb.seg = seg;
    a.rb.vis = vis;
    uint32_t* cnt = film->rb.counts;
    hipLaunchKernelGGL(k_camera, blocks_for(npaths), dim3(256), 0, s, a);
    uint32_t cur = 0;
    for (uint32_t k = 0; k < cmax && e == hipSuccess; ++k) {
        e = launch_intersect(render_setup(sc), reinterpret_cast<const gr_ray*>(film->rb.rays[cur]), (uint32_t)npaths, cnt + cur,
                             reinterpret_cast<gr_hit*>(film->rb.hits), nullptr, s);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, cnt, cur ^ 1u);
        hipLaunchKernelGGL(k_walk<true>, blocks_for(npaths), dim3(256), 0, s, a, b, k, cur);
        cur ^= 1u;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, cnt, 0u);
    hipLaunchKernelGGL(k_light_start, blocks_for(npaths), dim3(256), 0, s, a, b);
    cur = 0;
    for (uint32_t k = 1; k < lmax && e == hipSuccess; ++k) {
        e = launch_intersect(render_setup(sc), reinterpret_cast<const gr_ray*>(film->rb.rays[cur]), (uint32_t)npaths, cnt + cur,
                             reinterpret_cast<gr_hit*>(film->rb.hits), nullptr, s);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, cnt, cur ^ 1u);
        hipLaunchKernelGGL(k_walk<false>, blocks_for(npaths), dim3(256), 0, s, a, b, k, cur);
        cur ^= 1u;
    }
    for (size_t p0 = 0; p0 < npaths && e == hipSuccess; p0 += chunk) {
        const size_t p1 = std::min(npaths, p0 + chunk);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, cnt, 2u);
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, cnt, 4u);
        if (mis == 2) hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, cnt, 5u);
        hipLaunchKernelGGL(k_connect_list, blocks_for(p1 - p0), dim3(256), 0, s, a, b, (uint32_t)p0, (uint32_t)p1);
        // grid-stride over the list: enough waves to fill the chip, never more than the list can need
        const dim3 eval_grid(std::min<unsigned>(blocks_for((p1 - p0) * nstrat).x, 8192u));
        hipLaunchKernelGGL(k_connect_eval<false>, eval_grid, dim3(256), 0, s, a, b, (uint32_t)p0, (uint32_t)p1);
        if (mis == 2)
            hipLaunchKernelGGL(k_connect_eval<true>, dim3(std::min<unsigned>(blocks_for((p1 - p0) * lmax).x, 8192u)),
                               dim3(256), 0, s, a, b, (uint32_t)p0, (uint32_t)p1);
        e = launch_occluded(render_setup(sc), reinterpret_cast<const gr_segment*>(seg), (uint32_t)shadow, cnt + 2,
                            vis, GR_OCCLUSION_EXACT, nullptr, s);
        hipLaunchKernelGGL(k_vis_scatter, blocks_for(shadow), dim3(256), 0, s, a, b);
        hipLaunchKernelGGL(k_connect_sum, blocks_for(p1 - p0), dim3(256), 0, s, a, b, (uint32_t)p0, (uint32_t)p1);
        if (mis == 2)
            hipLaunchKernelGGL(k_lt_splat, blocks_for(p1 - p0), dim3(256), 0, s, a, b, (uint32_t)p0, (uint32_t)p1);
    }
    if (e != hipSuccess) return e;
    if (sample_out) hipLaunchKernelGGL(k_export_samples, blocks_for(npaths), dim3(256), 0, s, a, sample_out);
    const dim3 gather_grid((a.width + kTile - 1) / kTile, (a.height + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_film_gather, gather_grid, dim3(kTile * kTile), 0, s, a, film->d_pixels, nb,
                       b.mis == 2 ? b.splat : nullptr, b.mis == 2 ? b.ew : nullptr);
    return hipGetLastError();
}
}  // namespace

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
namespace {
// Guided-trace buffers of a film: per path a query (raw inputs, mixture) and
// a slot, plus the training vertex records (zeroed once; each batch's
// finalize clears what it read).
// The learned lobe choice's sums, counts and probabilities for scene sc,
// zeroed (probabilities unset) when the guide comes to a scene other than the
// one they were gathered on.
hipError_t reserve_lobe_table(gr_guide* g, const gr_scene* sc, hipStream_t s) {
    const size_t np = sc->prims.size();
    if (g->lobe_scene == sc && g->lobe_prims == np) return hipSuccess;
    const size_t cells = np * kLobeCells;
    const size_t b_sum = ((cells * 4 * 8) + 255) & ~(size_t)255, b_cnt = ((cells * 4 * 4) + 255) & ~(size_t)255;
    const size_t b_c = cells * sizeof(float4) + cells * sizeof(float);
    if (np > g->lobe_prims || !g->lobe_blob) {
        (void)hipFree(g->lobe_blob);
        g->lobe_blob = nullptr;
        g->lobe_prims = 0;
        g->lobe_scene = nullptr;
        hipError_t e = hipMalloc(&g->lobe_blob, b_sum + b_cnt + b_c);
        if (e != hipSuccess) return e;
    }
    char* b = static_cast<char*>(g->lobe_blob);
    g->lobe_sum = reinterpret_cast<unsigned long long*>(b);
    g->lobe_cnt = reinterpret_cast<uint32_t*>(b + b_sum);
    g->lobe_c = reinterpret_cast<float4*>(b + b_sum + b_cnt);
    g->lobe_m = reinterpret_cast<float*>(b + b_sum + b_cnt + cells * sizeof(float4));
    hipError_t e = hipMemsetAsync(b, 0, b_sum + b_cnt, s);
    if (e == hipSuccess) e = hipMemsetAsync(g->lobe_c, 0xff, b_c, s);  // NaN: no probabilities yet
    if (e != hipSuccess) return e;
    g->lobe_prims = np;
    g->lobe_scene = sc;
    return hipSuccess;
}

hipError_t reserve_guide(RenderBuffers& rb, size_t paths, size_t vrecs) {
    if (rb.guide_capacity >= paths && rb.guide_vrecs >= vrecs) return hipSuccess;
    paths = std::max(paths, rb.guide_capacity);
    vrecs = std::max(vrecs, rb.guide_vrecs);
    (void)hipFree(rb.guide_blob);
    rb.guide_blob = nullptr;
    rb.guide_capacity = rb.guide_vrecs = 0;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_raw = off; off = align(off + paths * 8 * 4);
    const size_t o_mix = off; off = align(off + paths * kMixFloats * 4);
    const size_t o_slot = off; off = align(off + paths * 4);
    const size_t o_vrec = off; off = align(off + vrecs * kVrecFloat4 * 16);
    char* b = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&b), off);
    if (e != hipSuccess) return e;
    e = hipMemset(b + o_vrec, 0, vrecs * kVrecFloat4 * 16);
    if (e != hipSuccess) {
        (void)hipFree(b);
        return e;
    }
    rb.guide_blob = b;
    rb.guide_capacity = paths;
    rb.guide_vrecs = vrecs;
    rb.graw = reinterpret_cast<float*>(b + o_raw);
    rb.gmix = reinterpret_cast<float*>(b + o_mix);
    rb.gslot = reinterpret_cast<uint32_t*>(b + o_slot);
    rb.vrec = reinterpret_cast<float4*>(b + o_vrec);
    return hipSuccess;
}
}  // namespace

hipError_t render_trace(gr_scene* sc, const gr_camera& cam, const gr_trace_params& prm, gr_film* film) {
    hipStream_t s = sc->ctx->stream;
    const uint32_t W = film->width, H = film->height;
    const size_t npx = (size_t)W * H;
    const bool guided = prm.integrator == GR_INTEGRATOR_GUIDED;
    gr_guide* guide = guided ? prm.guide : nullptr;
    if (guided && !guide) return hipErrorInvalidValue;
    uint32_t batch = prm.batch;
    if (batch == 0)
        batch = (uint32_t)std::max<size_t>(1, std::min<size_t>(prm.spp, (1u << 24) / npx));
    batch = std::min(batch, prm.spp);
    const size_t paths = npx * batch;
    if (paths > 0x7fffffffull) return hipErrorInvalidValue;
    hipError_t e = reserve(film->rb, paths);
    if (e == hipSuccess && sc->tuning.sort_bits) e = reserve_sort(sc->sort, paths);
    if (e != hipSuccess) return e;

    RenderArgs a;
    a.sv = sc->view;
    a.sh = sc->shade;
    a.width = W;
    a.height = H;
    a.max_depth = prm.max_depth;
    a.seed = prm.seed;
    camera_matrix(cam, a.cam);
    a.wx = cam.horizontal_aperture / (cam.focal_length * 2.0f);  // window_max, camera.rs:76-79
    a.wy = cam.vertical_aperture / (cam.focal_length * 2.0f);
    a.rb = film->rb;
    a.stats = sc->ctx->d_stats;
    a.ring_base = a.ring_cap = 0;
    std::memset(&a.gd, 0, sizeof(a.gd));
    const bool nee = prm.integrator == GR_INTEGRATOR_PT_NEE || guided;
    uint32_t nslots = 0;
    if (guided) {
        const gr_guide_params& gp = guide->prm;
        const uint32_t drec = prm.max_depth == 0 ? 16u : std::min(prm.max_depth, 16u);
        // a batch has at most `paths` training paths, whatever the guide asks for
        const uint64_t train = std::min<uint64_t>(gp.train_paths, paths);
        if (train * drec > 0x7fffffffull) return hipErrorInvalidValue;
        nslots = (uint32_t)(train * drec);
        e = reserve_guide(film->rb, paths, nslots);
        if (e == hipSuccess) e = guide_reserve_records(guide, nslots);
        if (e != hipSuccess) return e;
        a.rb = film->rb;
        a.gd.raw = film->rb.graw;
        a.gd.mix = film->rb.gmix;
        a.gd.gslot = film->rb.gslot;
        a.gd.vrec = film->rb.vrec;
        a.gd.recs = guide->recs;
        a.gd.rec_count = guide->count;
        a.gd.rec_cap = (uint32_t)guide->recs_cap;
        a.gd.train_paths = (uint32_t)train;
        a.gd.drec = drec;
        a.gd.guided_depth = gp.guided_depth;
        a.gd.lobes = gp.lobes;
        a.gd.select = gp.lobe_select;
        a.gd.select_mix = gp.select_mix;
        a.gd.rr = gp.rr_threshold;
        if (gp.lobe_select == GR_GUIDE_SELECT_LEARNED) {
            e = reserve_lobe_table(guide, sc, s);
            if (e != hipSuccess) return e;
            a.gd.lobe_sum = guide->lobe_sum;
            a.gd.lobe_cnt = guide->lobe_cnt;
            a.gd.lobe_c = guide->lobe_c;
            a.gd.lobe_m = guide->lobe_m;
        }
        a.gd.alpha = gp.alpha;
        for (int k = 0; k < 3; ++k) {
            const float ext = sc->view.root_hi[k] - sc->view.root_lo[k];
            a.gd.lo[k] = sc->view.root_lo[k];
            a.gd.inv_ext[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
        }
    }
    if (prm.integrator == GR_INTEGRATOR_BDPT || prm.integrator == GR_INTEGRATOR_BDPT_MIS ||
        prm.integrator == GR_INTEGRATOR_BDPT_LT) {
        const uint32_t cmax = (prm.max_depth == 0 || prm.max_depth > kBdptMaxDepth) ? kBdptMaxDepth : prm.max_depth;
        const uint32_t lmax =
            (prm.light_depth == 0 || prm.light_depth + 1 > kBdptMaxDepth) ? kBdptMaxDepth : prm.light_depth + 1;
        // vertex storage bounded to 32 GiB per batch (1080p, depths 8 + 6: 8 spp,
        // 26 GB): larger batches keep the traversal and connection launches full
        const size_t per_iter = npx * (cmax + lmax) * sizeof(BVert);
        uint32_t bb = (uint32_t)std::max<size_t>(1, ((size_t)32 << 30) / per_iter);
        bb = std::min(bb, batch);
        for (uint32_t done = 0; done < prm.spp && e == hipSuccess; done += bb) {
            const uint32_t nb = std::min(bb, prm.spp - done);
            a.npaths = (uint32_t)(npx * nb);
            a.path_base = (prm.sample_offset + done) * (uint64_t)npx;
            float4* out = prm.d_sample_radiance ? reinterpret_cast<float4*>(prm.d_sample_radiance) + (size_t)done * npx
                                                : nullptr;
            const uint32_t mis = prm.integrator == GR_INTEGRATOR_BDPT_LT ? 2u : prm.integrator == GR_INTEGRATOR_BDPT_MIS ? 1u : 0u;
            e = bdpt_batch(sc, a, film, cmax, lmax, nb, mis, out, s);
        }
        return e;
    }
    const dim3 gather_grid((W + kTile - 1) / kTile, (H + kTile - 1) / kTile);

    for (uint32_t done = 0; done < prm.spp && e == hipSuccess; done += batch) {
        const uint32_t nb = std::min(batch, prm.spp - done);
        a.npaths = (uint32_t)(npx * nb);
        a.path_base = (prm.sample_offset + done) * (uint64_t)npx;
        if (guided) {
            a.gd.active = guide->steps >= guide->prm.warmup_steps ? 1u : 0u;
            a.gd.train_stride = std::max<uint32_t>(1u, a.npaths / std::max<uint32_t>(1u, guide->prm.train_paths));
        }
        hipLaunchKernelGGL(k_camera, blocks_for(a.npaths), dim3(256), 0, s, a);
        uint32_t cur = 0;
        if (prm.integrator == GR_INTEGRATOR_DISPLAY) {
            e = launch_intersect(render_setup(sc), reinterpret_cast<const gr_ray*>(film->rb.rays[0]), a.npaths,
                                 nullptr, reinterpret_cast<gr_hit*>(film->rb.hits), nullptr, s);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(k_display, blocks_for(a.npaths), dim3(256), 0, s, a);
        }
        for (uint32_t depth = 0; prm.integrator != GR_INTEGRATOR_DISPLAY && (prm.max_depth == 0 || depth < prm.max_depth);
             ++depth) {
            if (prm.max_depth == 0 && depth > 0 && depth % 8 == 0) {
                // Unbounded paths (the reference's loop): stop once no path is live,
                // checked every 8 bounces (a bounce with no live path launches
                // kernels that find their counts zero)
                uint32_t live = 0;
                e = hipMemcpyAsync(&live, film->rb.counts + cur, 4, hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess || live == 0) break;
            }
            e = launch_intersect(render_setup(sc), reinterpret_cast<const gr_ray*>(film->rb.rays[cur]), a.npaths, film->rb.counts + cur,
                             reinterpret_cast<gr_hit*>(film->rb.hits), nullptr, s);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, film->rb.counts, cur ^ 1u);
            if (nee) {
                hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, film->rb.counts, 2u);
                if (guided && (a.gd.guided_depth == 0 || depth < a.gd.guided_depth)) {
                    hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, film->rb.counts, 3u);
                    hipLaunchKernelGGL(k_guide_query, blocks_for(a.npaths), dim3(256), 0, s, a, depth, cur);
                    if (a.gd.active) e = guide_infer(guide, a.gd.raw, a.npaths, film->rb.counts + 3, a.gd.mix, s);
                    if (e != hipSuccess) break;
                    hipLaunchKernelGGL((k_shade<true, true, true>), blocks_for(a.npaths), dim3(256), 0, s, a, depth,
                                       cur);
                } else if (guided) {
                    // beyond the guided depth: the lobe choice only, no queries
                    hipLaunchKernelGGL((k_shade<true, false, true>), blocks_for(a.npaths), dim3(256), 0, s, a, depth,
                                       cur);
                } else {
                    hipLaunchKernelGGL((k_shade<true, false, false>), blocks_for(a.npaths), dim3(256), 0, s, a, depth, cur);
                }
                e = launch_occluded(render_setup(sc), reinterpret_cast<const gr_segment*>(film->rb.seg), a.npaths,
                                    film->rb.counts + 2, film->rb.vis, GR_OCCLUSION_EXACT, nullptr, s);
                if (e != hipSuccess) break;
                if (guided)
                    hipLaunchKernelGGL((k_shadow_accum<true>), blocks_for(a.npaths), dim3(256), 0, s, a, depth);
                else
                    hipLaunchKernelGGL((k_shadow_accum<false>), blocks_for(a.npaths), dim3(256), 0, s, a, depth);
            } else {
                hipLaunchKernelGGL((k_shade<false, false, false>), blocks_for(a.npaths), dim3(256), 0, s, a, depth, cur);
            }
            e = hipGetLastError();
            cur ^= 1u;
        }
        if (e != hipSuccess) break;
        if (guided) {
            // train on this batch's records before the next batch samples the guide
            const dim3 rb = blocks_for(nslots);
            if (a.gd.select == GR_GUIDE_SELECT_LEARNED) {
                const uint32_t cells = (uint32_t)(guide->lobe_prims * kLobeCells);
                hipLaunchKernelGGL(k_lobe_stats, rb, dim3(256), 0, s, a, nslots, (uint32_t)guide->lobe_prims);
                hipLaunchKernelGGL(k_lobe_table, blocks_for(cells), dim3(256), 0, s, guide->lobe_sum, guide->lobe_cnt,
                                   sc->shade.mats, cells, a.gd.select_mix, guide->lobe_c, guide->lobe_m);
            }
            e = guide_zero_stats(guide, s);
            if (e != hipSuccess) break;
            hipLaunchKernelGGL(k_guide_rec_count, rb, dim3(256), 0, s, a, nslots, guide->rec_bcount);
            hipLaunchKernelGGL(k_guide_rec_scan, dim3(1), dim3(1024), 0, s, guide->rec_bcount, rb.x,
                               guide->rec_boff, guide->count);
            hipLaunchKernelGGL(k_guide_rec_write, rb, dim3(256), 0, s, a, nslots, guide->rec_boff);
            e = guide_record_stats(guide, guide->recs, (uint32_t)guide->recs_cap, guide->count, s);
            if (e == hipSuccess)
                e = guide_adam_steps(guide, guide->recs, (uint32_t)guide->recs_cap, guide->count,
                                     guide->prm.steps_per_pass, s);
            if (e != hipSuccess) break;
        }
        if (prm.d_sample_radiance)
            hipLaunchKernelGGL(k_export_samples, blocks_for(a.npaths), dim3(256), 0, s, a,
                               reinterpret_cast<float4*>(prm.d_sample_radiance) + (size_t)done * npx);
        hipLaunchKernelGGL(k_film_gather, gather_grid, dim3(kTile * kTile), 0, s, a, film->d_pixels, nb, nullptr,
                           nullptr);
        e = hipGetLastError();
    }
    return e;
}

hipError_t material_query(gr_scene* sc, uint32_t prim, const float* in, uint32_t n, float* out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_material_query, blocks_for(n), dim3(256), 0, sc->ctx->stream, sc->shade, prim, in, n, out);
    return hipGetLastError();
}

void release_render_buffers(RenderBuffers& rb) {
    (void)hipFree(rb.blob);
    (void)hipFree(rb.bdpt_blob);
    (void)hipFree(rb.guide_blob);
    rb = RenderBuffers{};
}


// ---------------------------------------------------------------------------
// Batched rendering: many jobs through one pool of in-flight paths
// ---------------------------------------------------------------------------
namespace {

struct JobDev {
    SceneView sv;
    ShadeView sh;
    float cam[16];
    float wx, wy;
    uint32_t width, height, max_depth, scene;
    uint64_t seed, path_base0;  // path id of the job's sample 0: sample_offset * width * height
    uint64_t first;             // index of the job's first sample among all jobs' samples
    uint32_t samples;
    uint32_t iter0;             // index of the job's first iteration among all jobs' iterations
};

struct BatchDev {
    const JobDev* jobs;
    uint32_t njobs, pool;
    unsigned long long total;   // samples of all jobs
    // Path state by queue position, one copy per queue side: shade_path's
    // throughput, radiance and MIS state, the path's job, sample within the job
    // and depth (x, y, z), and its film position.  Each bounce moves the
    // survivors' state to their positions in the next queue, so every access
    // is by queue position (coalesced) and a finished path's position is free
    // again at once.
    float4* T[2];
    float4* L[2];
    float4* aux[2];
    uint4* meta[2];
    float2* st[2];
    const unsigned long long* job_first;  // per job: index of its first sample (a compact search table)
    uint32_t* retire;    // queue positions (current side) of the paths that finished this step
    // [1] paths started this step, [3] their first queue position, [4] paths retired this step
    uint32_t* ctr;
    unsigned long long* issued;  // [0] samples started so far, [1] first sample of this step
    // Finished samples wait in a ring indexed by their global sample number:
    // radiance and film position at slot (job first + sample) mod ring.
    float4* ring_L;
    float2* ring_st;
    unsigned long long ring;
    // Samples may start only below *limit (set by the host: the first sample
    // not yet gathered into its film, plus ring), so the ring never overruns.
    const unsigned long long* limit;
    uint32_t* done;              // per iteration of every job: kDoneSpread counters of finished samples
    uint32_t niters;
};
constexpr uint32_t kDoneSpread = 16;  // spreads the counter updates of one iteration over 16 addresses
constexpr uint32_t kNextSide = 0x80000000u;  // shadow-ray owner flag: the path's state is in the next queue

__global__ void k_batch_init(RenderArgs a, BatchDev b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < b.niters * kDoneSpread) b.done[i] = 0u;
    if (i == 0) {
        b.ctr[1] = b.ctr[3] = b.ctr[4] = 0u;
        b.issued[0] = b.issued[1] = 0ull;
        a.rb.counts[0] = a.rb.counts[1] = a.rb.counts[2] = 0u;
    }
}

// How many new paths this step starts (room in the queue after the paths
// that continue, samples left) and where they go; clears the step's other
// queues.
__global__ void k_batch_plan(RenderArgs a, BatchDev b, uint32_t cur) {
    if (threadIdx.x != 0) return;
    const uint32_t q = a.rb.counts[cur];  // continuing paths at queue positions [0, q)
    const unsigned long long issued = b.issued[0];
    const unsigned long long lim = *b.limit;
    unsigned long long rem = b.total - issued;
    const unsigned long long avail = lim > issued ? lim - issued : 0ull;
    rem = avail < rem ? avail : rem;
    const uint32_t room = b.pool - q;
    const uint32_t k = rem < (unsigned long long)room ? (uint32_t)rem : room;
    b.ctr[1] = k;
    b.ctr[3] = q;
    b.ctr[4] = 0u;
    b.issued[1] = issued;
    b.issued[0] = issued + k;
    a.rb.counts[cur] = q + k;
    a.rb.counts[cur ^ 1u] = 0u;
    a.rb.counts[2] = 0u;
}

// New camera paths appended after the continuing ones.
__global__ __launch_bounds__(256) void k_batch_regen(RenderArgs a, BatchDev b, uint32_t cur) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.ctr[1]) return;
    const unsigned long long sidx = b.issued[1] + i;
    uint32_t lo = 0, hi = b.njobs - 1;  // last job whose first sample <= sidx
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (b.job_first[mid] <= sidx) lo = mid;
        else hi = mid - 1;
    }
    const JobDev& J = b.jobs[lo];
    const uint32_t local = (uint32_t)(sidx - J.first);
    const uint32_t px = local % (J.width * J.height);
    const CameraSample c =
        camera_sample(J.cam, J.wx, J.wy, J.width, J.height, J.seed, J.path_base0 + local, px / J.width, px % J.width);
    const uint32_t q = b.ctr[3] + i;
    a.rb.rays[cur][2 * (size_t)q] = make_float4(c.o.x, c.o.y, c.o.z, __uint_as_float(J.scene));
    a.rb.rays[cur][2 * (size_t)q + 1] = make_float4(c.d.x, c.d.y, c.d.z, 0.0f);
    b.T[cur][q] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    b.L[cur][q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    b.aux[cur][q] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);  // previous bsdf pdf, previous bounce specular
    b.meta[cur][q] = make_uint4(lo, local, 0u, 0u);
    b.st[cur][q] = c.st;
}

// One bounce of every queued path, each with its own job's scene, seed and
// depth; the shading itself is shade_path, as in the per-film wavefront (a's
// T / L / aux are this side's arrays, indexed by queue position).  Survivors
// move with their state to the next queue; finished paths are listed for
// retirement; a shadow ray records where its path's radiance now lives.
template <bool NEE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_shade(RenderArgs a, BatchDev b, uint32_t cur) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = a.rb.counts[cur];
    const uint32_t nxt = cur ^ 1u;
    ShadeOut out;
    out.alive = out.shadow = false;
    uint32_t scene = 0;
    uint4 m = make_uint4(0u, 0u, 0u, 0u);
    bool retire = false;
    if (i < n) {
        const uint4 hv = a.rb.hits[i];
        const gr_hit h = {hv.x, hv.y, __uint_as_float(hv.z), hv.w};
        const float4 ro = a.rb.rays[cur][2 * (size_t)i];
        const vec3 o = ld3(ro), d = ld3(a.rb.rays[cur][2 * (size_t)i + 1]);
        scene = __float_as_uint(ro.w);
        m = b.meta[cur][i];
        const JobDev& J = b.jobs[m.x];
        shade_path<NEE, false, false>(a, J.sv, J.sh, J.seed, J.max_depth, J.path_base0 + m.y, i, m.z, h, o, d, out);
        retire = !out.alive;
    }
    const uint32_t q = block_append(out.alive, &a.rb.counts[nxt]);
    if (out.alive) {
        a.rb.rays[nxt][2 * (size_t)q] = make_float4(out.new_o.x, out.new_o.y, out.new_o.z, __uint_as_float(scene));
        a.rb.rays[nxt][2 * (size_t)q + 1] = make_float4(out.new_d.x, out.new_d.y, out.new_d.z, 0.0f);
        b.T[nxt][q] = b.T[cur][i];
        b.L[nxt][q] = b.L[cur][i];
        b.aux[nxt][q] = b.aux[cur][i];
        b.meta[nxt][q] = make_uint4(m.x, m.y, m.z + 1u, 0u);
        b.st[nxt][q] = b.st[cur][i];
    }
    const uint32_t r = block_append(retire, &b.ctr[4]);
    if (retire) b.retire[r] = i;
    if (NEE) {
        const uint32_t ss = block_append(out.shadow, &a.rb.counts[2]);
        if (out.shadow) {
            a.rb.seg[2 * (size_t)ss] = make_float4(out.hp.x, out.hp.y, out.hp.z, __uint_as_float(scene));
            a.rb.seg[2 * (size_t)ss + 1] = make_float4(out.seg_to.x, out.seg_to.y, out.seg_to.z, 0.0f);
            a.rb.contrib[ss] = make_float4(out.c_nee.x, out.c_nee.y, out.c_nee.z, 0.0f);
            a.rb.seg_path[ss] = out.alive ? (q | kNextSide) : i;
        }
    }
}

// The light samples that reached their lights, added to their paths'
// radiance (in the next queue for survivors, in this one for finished paths).
__global__ __launch_bounds__(256) void k_batch_shadow_accum(RenderArgs a, BatchDev b, uint32_t cur) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.rb.counts[2] || a.rb.vis[j] == 0) return;
    const uint32_t owner = a.rb.seg_path[j];
    float4* Ls = (owner & kNextSide) ? b.L[cur ^ 1u] : b.L[cur];
    const uint32_t p = owner & ~kNextSide;
    const float4 c = a.rb.contrib[j];
    float4 L = Ls[p];
    L.x = L.x + c.x;
    L.y = L.y + c.y;
    L.z = L.z + c.z;
    Ls[p] = L;
}

// Finished paths (their last light sample already added): radiance and film
// position to the sample ring, one count to their iteration.
__global__ __launch_bounds__(256) void k_batch_retire(RenderArgs a, BatchDev b, uint32_t cur) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = i < b.ctr[4];
    uint32_t it = kNone;
    if (ok) {
        const uint32_t p = b.retire[i];
        const uint4 m = b.meta[cur][p];
        const JobDev& J = b.jobs[m.x];
        const size_t slot = (size_t)((J.first + m.y) % b.ring);
        b.ring_L[slot] = b.L[cur][p];
        b.ring_st[slot] = b.st[cur][p];
        it = J.iter0 + m.y / (J.width * J.height);
    }
    // one counter update per distinct iteration in the wave (finished paths come
    // roughly in sample order, so a wave usually holds one or two iterations)
    uint64_t left = __ballot(ok);
    while (left) {
        const uint32_t lead = (uint32_t)__builtin_ctzll(left);
        const uint32_t li = __shfl(it, (int)lead);
        const uint64_t same = __ballot(ok && it == li);
        if ((threadIdx.x & 63u) == lead)
            atomicAdd(&b.done[(size_t)li * kDoneSpread + blockIdx.x % kDoneSpread], (uint32_t)__popcll(same));
        left &= ~same;
    }
}

}  // namespace

hipError_t render_jobs(gr_ctx* ctx, const gr_job* jobs, uint32_t njobs, const gr_batch_params& prm, double* job_ms) {
    const bool nee = prm.integrator == GR_INTEGRATOR_PT_NEE;
    const uint32_t pool = prm.path_pool ? prm.path_pool : (1u << 23);
    const uint32_t poll = prm.steps_per_poll ? prm.steps_per_poll : 16u;
    // generation order: the jobs as given, or fewest samples first (stable)
    std::vector<uint32_t> order(njobs);
    for (uint32_t j = 0; j < njobs; ++j) order[j] = j;
    auto job_samples = [&](uint32_t j) { return (uint64_t)jobs[j].film->width * jobs[j].film->height * jobs[j].spp; };
    if (prm.job_order == GR_JOBS_FEWEST_FIRST)
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t x, uint32_t y) { return job_samples(x) < job_samples(y); });
    // distinct scenes, job table (in generation order: entry g is job order[g])
    std::vector<gr_scene*> scenes;
    std::vector<JobDev> hj(njobs);
    uint64_t total = 0, max_npx = 0;
    uint32_t niters = 0;
    for (uint32_t g = 0; g < njobs; ++g) {
        const gr_job& jb = jobs[order[g]];
        uint32_t si = 0;
        while (si < scenes.size() && scenes[si] != jb.scene) ++si;
        if (si == scenes.size()) scenes.push_back(jb.scene);
        JobDev& d = hj[g];
        std::memset(&d, 0, sizeof(d));
        d.sv = jb.scene->view;
        d.sh = jb.scene->shade;
        camera_matrix(jb.camera, d.cam);
        d.wx = jb.camera.horizontal_aperture / (jb.camera.focal_length * 2.0f);  // window_max, camera.rs:76-79
        d.wy = jb.camera.vertical_aperture / (jb.camera.focal_length * 2.0f);
        d.width = jb.film->width;
        d.height = jb.film->height;
        d.max_depth = jb.max_depth;
        d.scene = si;
        d.seed = jb.seed;
        const uint64_t npx = (uint64_t)d.width * d.height;
        max_npx = std::max(max_npx, npx);
        d.path_base0 = jb.sample_offset * npx;
        d.first = total;
        d.samples = (uint32_t)(npx * jb.spp);
        d.iter0 = niters;
        total += d.samples;
        niters += jb.spp;
    }
    // The sample ring: at least one whole iteration of the largest film, so the
    // oldest ungathered iteration always fits below the generation limit.  The
    // limit moves only at polls, and a poll interval generates about
    // steps_per_poll * pool / (mean path length) samples, so the default holds
    // 16 pools' worth (no more than the batch's samples).
    uint64_t ring = prm.sample_ring ? prm.sample_ring : std::min<uint64_t>(16ull * pool, total);
    ring = std::max(ring, max_npx);
    std::vector<SceneRef> hs(scenes.size());
    for (size_t k = 0; k < scenes.size(); ++k) {
        const SceneView& v = scenes[k]->view;
        hs[k].nodes = v.nodes;
        hs[k].comps = v.comps;
        hs[k].comp_ids = v.comp_ids;
        hs[k].root_ref = v.root_ref;
        hs[k].simple_boxes = v.simple_boxes;
        for (int c = 0; c < 3; ++c) {
            hs[k].root_lo[c] = v.root_lo[c];
            hs[k].root_hi[c] = v.root_hi[c];
        }
    }
    TraceTuning tu = scenes[0]->render_tuning;
    if (prm.tuning.lds_stack != 0) {
        tu.sort_bits = sort_key_bits(prm.tuning.sort_bits);
        tu.lds_stack = prm.tuning.lds_stack;
        tu.waves_per_cu = prm.tuning.waves_per_cu;
        tu.refill = prm.tuning.refill;
        tu.vis_refill = prm.tuning.vis_refill ? prm.tuning.vis_refill : prm.tuning.refill;
        tu.xcd_ranges = prm.tuning.xcd_ranges;
        tu.ray_chunk = prm.tuning.ray_chunk ? prm.tuning.ray_chunk : kDefaultRayChunk;
    }

    // device memory: job and scene tables, the sample ring, path state, counters
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_jobs = off; off = align(off + njobs * sizeof(JobDev));
    const size_t o_scenes = off; off = align(off + scenes.size() * sizeof(SceneRef));
    const size_t o_L = off; off = align(off + ring * sizeof(float4));
    const size_t o_st = off; off = align(off + ring * sizeof(float2));
    // the second queue side's path state (the first side uses the film buffers' T, L, aux, st)
    const size_t o_T1 = off; off = align(off + (size_t)pool * 16);
    const size_t o_L1 = off; off = align(off + (size_t)pool * 16);
    const size_t o_aux1 = off; off = align(off + (size_t)pool * 16);
    const size_t o_st1 = off; off = align(off + (size_t)pool * 8);
    const size_t o_meta0 = off; off = align(off + (size_t)pool * 16);
    const size_t o_meta1 = off; off = align(off + (size_t)pool * 16);
    const size_t o_first = off; off = align(off + (size_t)njobs * 8);
    const size_t o_ret = off; off = align(off + (size_t)pool * 4);
    const size_t o_ctr = off; off = align(off + 64);
    const size_t o_iss = off; off = align(off + 16);
    const size_t o_lim = off; off = align(off + 8);
    const size_t o_done = off; off = align(off + (size_t)niters * kDoneSpread * 4);
    char* blob = nullptr;
    RenderBuffers rb;
    SortScratch ss;
    // Polls are double-buffered: the steps of the next interval are queued
    // before the host reads the previous interval's counters, and film gathers
    // run on a stream of their own, so the GPU never waits for the host and
    // gathers overlap the bounce steps.
    uint32_t* h_done[2] = {nullptr, nullptr};
    unsigned long long* h_limit = nullptr;  // two words, used alternately
    hipStream_t s = nullptr, sg = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev_start = nullptr, ev_ctx = nullptr, ev_poll[2] = {nullptr, nullptr}, ev_gath[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> ev_job(njobs, nullptr);  // by generation index
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&blob), off);
    if (e == hipSuccess) e = reserve(rb, pool);
    if (e == hipSuccess && tu.sort_bits) e = reserve_sort(ss, pool);
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
        e = hipHostMalloc(reinterpret_cast<void**>(&h_done[k]), (size_t)niters * kDoneSpread * 4);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&h_limit), 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sg, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ev_start);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipEventCreateWithFlags(&ev_poll[k], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_gath[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_ctx, hipEventDisableTiming);
    for (uint32_t j = 0; j < njobs && e == hipSuccess; ++j) e = hipEventCreate(&ev_job[j]);
    std::vector<unsigned long long> job_first(njobs);
    for (uint32_t g = 0; g < njobs; ++g) job_first[g] = hj[g].first;
    if (e == hipSuccess) {
        h_limit[0] = h_limit[1] = ring;  // nothing gathered yet
        // start after the work already on the context's stream
        e = hipEventRecord(ev_ctx, ctx->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, ev_ctx, 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(blob + o_jobs, hj.data(), njobs * sizeof(JobDev), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(blob + o_scenes, hs.data(), hs.size() * sizeof(SceneRef), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(blob + o_first, job_first.data(), njobs * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(blob + o_lim, h_limit, 8, hipMemcpyHostToDevice, s);
    }
    RenderArgs a;
    std::memset(&a, 0, sizeof(a));
    a.rb = rb;
    a.stats = ctx->d_stats;
    BatchDev b;
    b.jobs = reinterpret_cast<const JobDev*>(blob + o_jobs);
    b.njobs = njobs;
    b.pool = pool;
    b.total = total;
    b.T[0] = rb.T;
    b.L[0] = rb.L;
    b.aux[0] = rb.aux;
    b.st[0] = rb.st;
    b.T[1] = reinterpret_cast<float4*>(blob + o_T1);
    b.L[1] = reinterpret_cast<float4*>(blob + o_L1);
    b.aux[1] = reinterpret_cast<float4*>(blob + o_aux1);
    b.st[1] = reinterpret_cast<float2*>(blob + o_st1);
    b.meta[0] = reinterpret_cast<uint4*>(blob + o_meta0);
    b.meta[1] = reinterpret_cast<uint4*>(blob + o_meta1);
    b.job_first = reinterpret_cast<const unsigned long long*>(blob + o_first);
    b.retire = reinterpret_cast<uint32_t*>(blob + o_ret);
    b.ctr = reinterpret_cast<uint32_t*>(blob + o_ctr);
    b.issued = reinterpret_cast<unsigned long long*>(blob + o_iss);
    b.ring_L = reinterpret_cast<float4*>(blob + o_L);
    b.ring_st = reinterpret_cast<float2*>(blob + o_st);
    b.ring = ring;
    b.limit = reinterpret_cast<const unsigned long long*>(blob + o_lim);
    b.done = reinterpret_cast<uint32_t*>(blob + o_done);
    b.niters = niters;

    TraceSetup ts{scenes[0]->view, tu, &ss, nullptr, reinterpret_cast<const SceneRef*>(blob + o_scenes),
                  (uint32_t)scenes.size()};
    auto step = [&](uint32_t cur) -> hipError_t {
        RenderArgs ac = a;  // shade_path's state arrays: this queue side's
        ac.rb.T = b.T[cur];
        ac.rb.L = b.L[cur];
        ac.rb.aux = b.aux[cur];
        hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(64), 0, s, a, b, cur);
        hipLaunchKernelGGL(k_batch_regen, blocks_for(pool), dim3(256), 0, s, a, b, cur);
        hipError_t r = launch_intersect(ts, reinterpret_cast<const gr_ray*>(rb.rays[cur]), pool, rb.counts + cur,
                                        reinterpret_cast<gr_hit*>(rb.hits), nullptr, s);
        if (r != hipSuccess) return r;
        if (nee) {
            hipLaunchKernelGGL((k_batch_shade<true>), blocks_for(pool), dim3(256), 0, s, ac, b, cur);
            r = launch_occluded(ts, reinterpret_cast<const gr_segment*>(rb.seg), pool, rb.counts + 2, rb.vis,
                                GR_OCCLUSION_EXACT, nullptr, s);
            if (r != hipSuccess) return r;
            hipLaunchKernelGGL(k_batch_shadow_accum, blocks_for(pool), dim3(256), 0, s, a, b, cur);
        } else {
            hipLaunchKernelGGL((k_batch_shade<false>), blocks_for(pool), dim3(256), 0, s, ac, b, cur);
        }
        hipLaunchKernelGGL(k_batch_retire, blocks_for(pool), dim3(256), 0, s, a, b, cur);
        return hipGetLastError();
    };
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_batch_init, blocks_for(std::max(1u, niters * kDoneSpread)), dim3(256), 0, s, a, b);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev_start, s);
    // two eager steps (they also settle the launch shapes), then the graph
    if (e == hipSuccess) e = step(0);
    if (e == hipSuccess) e = step(1);
    if (e == hipSuccess && prm.use_graph) {
        e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t e1 = step(0);
            if (e1 == hipSuccess) e1 = step(1);
            e = hipStreamEndCapture(s, &graph);
            if (e == hipSuccess) e = e1;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    }
    std::vector<uint32_t> gathered(njobs, 0);  // iterations gathered, by generation index
    uint32_t pending = njobs;
    uint64_t pairs = 0;
    // Every path ends: past 20 bounces Russian roulette keeps at most 3/4 of
    // them per bounce.  A workload needs about (samples / pool) * depth steps
    // (more when the ring holds generation back); a count far beyond that means
    // the counters are wrong -- stop with an error.
    const uint64_t max_pairs = (total / pool + 2) * 4096ull;
    // One interval of steps, then a copy of the iteration counters into
    // h_done[k] and ev_poll[k] behind it.
    auto interval = [&](int k) -> hipError_t {
        hipError_t r = hipSuccess;
        for (uint32_t i = 0; i < (poll + 1) / 2 && r == hipSuccess; ++i) {
            if (exec) {
                r = hipGraphLaunch(exec, s);
            } else {
                r = step(0);
                if (r == hipSuccess) r = step(1);
            }
            ++pairs;
        }
        if (r == hipSuccess)
            r = hipMemcpyAsync(h_done[k], b.done, (size_t)niters * kDoneSpread * 4, hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipEventRecord(ev_poll[k], s);
        return r;
    };
    // The limit moves to (first ungathered sample + ring) only once the gathers
    // that freed those ring slots are complete: one poll behind the gathers.
    uint64_t lo_issued = 0;  // first ungathered sample when the last gathers were queued
    bool gath_pending = false;
    int cur = 0;
    if (e == hipSuccess) e = interval(0);
    while (e == hipSuccess && pending > 0) {
        e = interval(cur ^ 1);  // the GPU keeps working while the host reads poll `cur`
        if (e == hipSuccess) e = hipEventSynchronize(ev_poll[cur]);
        if (e == hipSuccess && gath_pending) {
            // the previous poll's gathers have had a whole interval to finish
            e = hipEventSynchronize(ev_gath[cur ^ 1]);
            if (e == hipSuccess) {
                // ev_poll[cur] follows the last copy from this word, so it may be rewritten
                h_limit[cur] = lo_issued + ring;
                e = hipMemcpyAsync(blob + o_lim, h_limit + cur, 8, hipMemcpyHostToDevice, s);
            }
        }
        // each job's iterations whose samples are all in, in order, into its film
        // (on the gather stream, after the steps that wrote them)
        if (e == hipSuccess) e = hipStreamWaitEvent(sg, ev_poll[cur], 0);
        uint64_t lo = total;  // first sample not yet gathered
        const uint32_t* done = h_done[cur];
        for (uint32_t g = 0; g < njobs && e == hipSuccess; ++g) {
            const JobDev& d = hj[g];
            const gr_job& jb = jobs[order[g]];
            const uint32_t npx = d.width * d.height;
            uint32_t k1 = gathered[g];
            while (k1 < jb.spp) {
                uint64_t fin = 0;
                for (uint32_t c = 0; c < kDoneSpread; ++c) fin += done[(size_t)(d.iter0 + k1) * kDoneSpread + c];
                if (fin < npx) break;
                ++k1;
            }
            if (k1 > gathered[g]) {
                RenderArgs ga = a;
                ga.width = d.width;
                ga.height = d.height;
                ga.rb.L = b.ring_L;
                ga.rb.st = b.ring_st;
                ga.ring_base = d.first + (uint64_t)gathered[g] * npx;
                ga.ring_cap = ring;
                const dim3 grid((ga.width + kTile - 1) / kTile, (ga.height + kTile - 1) / kTile);
                hipLaunchKernelGGL(k_film_gather, grid, dim3(kTile * kTile), 0, sg, ga, jb.film->d_pixels,
                                   k1 - gathered[g], nullptr, nullptr);
                e = hipGetLastError();
                if (e == hipSuccess && k1 == jb.spp) {
                    e = hipEventRecord(ev_job[g], sg);
                    --pending;
                }
                gathered[g] = k1;
            }
            if (k1 < jb.spp) lo = std::min(lo, d.first + (uint64_t)k1 * npx);
        }
        if (e == hipSuccess) e = hipEventRecord(ev_gath[cur], sg);
        lo_issued = lo;
        gath_pending = true;
        cur ^= 1;
        if (e == hipSuccess && pairs > max_pairs) e = hipErrorLaunchTimeOut;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(sg);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (uint32_t g = 0; g < njobs && e == hipSuccess && job_ms; ++g) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, ev_start, ev_job[g]);
        job_ms[order[g]] = ms;
    }
    // the context's stream continues after the batch
    if (ev_ctx && s) {
        (void)hipEventRecord(ev_ctx, s);
        (void)hipStreamWaitEvent(ctx->stream, ev_ctx, 0);
    }
    if (ev_ctx && sg) {
        (void)hipEventRecord(ev_ctx, sg);
        (void)hipStreamWaitEvent(ctx->stream, ev_ctx, 0);
    }
    if (s) (void)hipStreamSynchronize(s);
    if (sg) (void)hipStreamSynchronize(sg);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t ev : ev_job)
        if (ev) (void)hipEventDestroy(ev);
    if (ev_start) (void)hipEventDestroy(ev_start);
    for (int k = 0; k < 2; ++k) {
        if (ev_poll[k]) (void)hipEventDestroy(ev_poll[k]);
        if (ev_gath[k]) (void)hipEventDestroy(ev_gath[k]);
    }
    if (ev_ctx) (void)hipEventDestroy(ev_ctx);
    if (s) (void)hipStreamDestroy(s);
    if (sg) (void)hipStreamDestroy(sg);
    for (int k = 0; k < 2; ++k)
        if (h_done[k]) (void)hipHostFree(h_done[k]);
    if (h_limit) (void)hipHostFree(h_limit);
    release_sort(ss);
    release_render_buffers(rb);
    (void)hipFree(blob);
    return e;
}

}  // namespace gr
