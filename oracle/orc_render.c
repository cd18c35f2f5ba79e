/*
 * orc_render.c -- CPU oracle: camera, film, light sampling and the
 * integrators (test infrastructure only; see gr_oracle.h).
 *
 * Integrators:
 *   ORC_PT      PathTracerIntegrator (src/render/integrators.rs:39-100) plus a
 *               max_depth cap on the number of intersections (0 = none).
 *   ORC_PT_NEE  path tracing with next-event estimation and power-heuristic
 *               MIS.  The reference has no such estimator; this is the
 *               specification the GPU kernels follow (parity unpinned by the
 *               reference; pinned here by agreement with ORC_PT's mean on
 *               pdf-consistent materials and by analytic scenes).
 *   ORC_BDPT    BdptIntegrator (integrators.rs:102-418) with caps on camera
 *               vertices (max_depth) and light vertices (1 + light_depth).
 *   ORC_DISPLAY DisplayColorIntegrator (integrators.rs:18-34).
 */
#include "gr_oracle.h"
#include "orc_render.h"

#include <stdlib.h>
#include <string.h>

#define FILTER_WIDTH 2.0f /* src/render/film.rs:8 */

/* ---- camera, src/core/camera.rs, matrix.rs, xform.rs ---------------------- */

static void mat_identity(float m[4][4])
{
    memset(m, 0, 16 * sizeof(float));
    m[0][0] = m[1][1] = m[2][2] = m[3][3] = 1.0f;
}

/* Mat * Mat, src/core/matrix.rs:317-400 (row-major, left to right sums). */
static void mat_mul(const float a[4][4], const float b[4][4], float out[4][4])
{
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            out[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j] + a[i][3] * b[3][j];
}

/* Camera::new (camera.rs:42-61): xform = translation * rotation, built by
 * Mat::translation (matrix.rs:73-93) and Mat::rotation (matrix.rs:95-121). */
void orc_camera_matrix(const float quat[4], const float translate[3], float out[16])
{
    float t[4][4], r[4][4], m[4][4];
    mat_identity(t);
    t[3][0] = translate[0]; t[3][1] = translate[1]; t[3][2] = translate[2];
    float rr = quat[0], ix = quat[1], iy = quat[2], iz = quat[3];
    memset(r, 0, sizeof(r));
    r[0][0] = 1.0f - 2.0f * (iy * iy + iz * iz);
    r[0][1] = 2.0f * (ix * iy + iz * rr);
    r[0][2] = 2.0f * (iz * ix - iy * rr);
    r[1][0] = 2.0f * (ix * iy - iz * rr);
    r[1][1] = 1.0f - 2.0f * (iz * iz + ix * ix);
    r[1][2] = 2.0f * (iy * iz + ix * rr);
    r[2][0] = 2.0f * (iz * ix + iy * rr);
    r[2][1] = 2.0f * (iy * iz - ix * rr);
    r[2][2] = 1.0f - 2.0f * (iy * iy + ix * ix);
    r[3][3] = 1.0f;
    mat_mul(t, r, m);
    memcpy(out, m, sizeof(m));
}

/* Camera::compute_ray (camera.rs:87-97) with Xform::_transform /
 * _transform_dir (xform.rs:41-54). */
static void camera_ray(const float *m, float wx, float wy, float s, float t, ovec *o, ovec *d)
{
    ovec dir = v_normalized(v_new(wx * s, wy * t, -1.0f));
    ovec z = v_zero();
    float x = z.x * m[0] + z.y * m[4] + z.z * m[8] + m[12];
    float y = z.x * m[1] + z.y * m[5] + z.z * m[9] + m[13];
    float zz = z.x * m[2] + z.y * m[6] + z.z * m[10] + m[14];
    float w = z.x * m[3] + z.y * m[7] + z.z * m[11] + m[15];
    *o = v_new(x / w, y / w, zz / w);
    *d = v_new(dir.x * m[0] + dir.y * m[4] + dir.z * m[8],
               dir.x * m[1] + dir.y * m[5] + dir.z * m[9],
               dir.x * m[2] + dir.y * m[6] + dir.z * m[10]);
}

void orc_camera_rays(const float cam[16], float wx, float wy, const float *st, int64_t n, float *rays)
{
    for (int64_t i = 0; i < n; ++i) {
        ovec o, d;
        camera_ray(cam, wx, wy, st[2 * i], st[2 * i + 1], &o, &d);
        rays[6*i] = o.x; rays[6*i+1] = o.y; rays[6*i+2] = o.z;
        rays[6*i+3] = d.x; rays[6*i+4] = d.y; rays[6*i+5] = d.z;
    }
}

/* ---- film, src/render/film.rs, src/core/math.rs:92-123 -------------------- */

float orc_mitchell1(float x)
{
    const float B = 1.0f / 3.0f;
    const float C = 1.0f / 3.0f;
    float twox = fabsf(2.0f * x);
    if (twox > 1.0f) {
        return ((-B - 6.0f * C) * (twox * twox * twox) + (6.0f * B + 30.0f * C) * (twox * twox) +
                (-12.0f * B - 48.0f * C) * twox + (8.0f * B + 24.0f * C)) * (1.0f / 6.0f);
    }
    return ((12.0f - 9.0f * B - 6.0f * C) * (twox * twox * twox) +
            (-18.0f + 12.0f * B + 6.0f * C) * (twox * twox) + (6.0f - 2.0f * B)) * (1.0f / 6.0f);
}

static float mitchell2(float x, float y) { return orc_mitchell1(x / FILTER_WIDTH) * orc_mitchell1(y / FILTER_WIDTH); }

/* Film::compute_sample_points (film.rs:55-76) for one pixel: the row jitter is
 * drawn before the column jitter; Range(-2, 2) -> -2 + 4 u. */
static void sample_point(uint64_t seed, uint64_t path, uint32_t row, uint32_t col, uint32_t w,
                         uint32_t h, float *s, float *t)
{
    float row_cont = 0.5f + (float)row;
    float col_cont = 0.5f + (float)col;
    float rj = row_cont + (-FILTER_WIDTH + 2.0f * FILTER_WIDTH * orc_draw_f32(seed, path, ORC_STREAM_FILM, 0, 0));
    float cj = col_cont + (-FILTER_WIDTH + 2.0f * FILTER_WIDTH * orc_draw_f32(seed, path, ORC_STREAM_FILM, 0, 1));
    *s = lerpf(-1.0f, 1.0f, cj / (float)w);
    *t = lerpf(-1.0f, 1.0f, rj / (float)h);
}

static int64_t clamp_i(int64_t x, int64_t a, int64_t b) { return x < a ? a : (x > b ? b : x); }

/* f32 -> isize as Rust casts it (saturating, NaN -> 0). */
static int64_t f2i(float f)
{
    if (f != f) return 0;
    if (f >= 9.2233720368547758e18f) return INT64_MAX;
    if (f <= -9.2233720368547758e18f) return INT64_MIN;
    return (int64_t)f;
}

/* Film::report_samples, film.rs:78-114 (serial, in sample order). */
static void report_samples(float *film, uint32_t w, uint32_t h, const float *st, const float *col,
                           uint64_t n)
{
    float wf = (float)w, hf = (float)h;
    int64_t last_col = (int64_t)w - 1, last_row = (int64_t)h - 1;
    for (uint64_t i = 0; i < n; ++i) {
        float s = st[2 * i], t = st[2 * i + 1];
        float col_cont = lerpf(0.0f, wf, 0.5f * (s + 1.0f));
        float row_cont = lerpf(0.0f, hf, 0.5f * (t + 1.0f));
        float cd = col_cont - 0.5f, rd = row_cont - 0.5f;
        int64_t min_col = clamp_i(f2i(ceilf(cd - FILTER_WIDTH)), 0, last_col);
        int64_t max_col = clamp_i(f2i(floorf(cd + FILTER_WIDTH)), 0, last_col);
        int64_t min_row = clamp_i(f2i(ceilf(rd - FILTER_WIDTH)), 0, last_row);
        int64_t max_row = clamp_i(f2i(floorf(rd + FILTER_WIDTH)), 0, last_row);
        for (int64_t y = min_row; y < max_row + 1; ++y) {
            for (int64_t x = min_col; x < max_col + 1; ++x) {
                float *px = &film[4 * (y * (int64_t)w + x)];
                float wt = mitchell2((float)x - cd, (float)y - rd);
                px[0] = px[0] + col[3 * i] * wt;
                px[1] = px[1] + col[3 * i + 1] * wt;
                px[2] = px[2] + col[3 * i + 2] * wt;
                px[3] += wt;
            }
        }
    }
}

/* ---- light sampling, src/geom/bvh.rs:481-496, prim.rs:37-51 --------------- */

typedef struct {
    ovec point, dir;
    orc_frame props;
    uint32_t prim;
    float point_pdf, dir_pdf;
} light_sample;

/* CumulativeDistribution::ind_sample (src/core/random.rs:248-255): index of
 * the first cdf entry >= target. */
static uint32_t cdf_search(const float *cdf, uint32_t n, float target)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (cdf[mid] < target) lo = mid + 1; else hi = mid;
    }
    return lo < n ? lo : n - 1;
}

/* Prim::sample_world for meshes (mesh.rs:310-327) and spheres
 * (sphere.rs:103-111). */
static void prim_sample_point(const orc_prim *p, float u_tri, float u1, float u2, ovec *pt,
                              orc_frame *props, float *pdf)
{
    if (p->type == ORC_PRIM_MESH) {
        uint32_t ti = cdf_search(p->area_cdf, p->ntris, u_tri);
        const uint32_t *t = &p->tris[9 * (size_t)ti];
        ovec a = p->verts[t[0]], b = p->verts[t[1]], c = p->verts[t[2]];
        float sa = sqrtf(u1);
        float u = 1.0f - sa, v = u2 * sa; /* UniformSampleBarycentric, random.rs:268-274 */
        float w = 1.0f - u - v;
        *pt = v_add(v_add(v_scale(a, u), v_scale(b, v)), v_scale(c, w));
        orc_mesh_frame(p, ti, u, v, w, props);
        *pdf = 1.0f / p->area;
    } else {
        *pt = v_add(p->center, v_scale(orc_uniform_sphere(u1, u2), p->radius));
        orc_sphere_frame(p, *pt, props);
        *pdf = 1.0f / (4.0f * ORC_PI * p->radius * p->radius);
    }
}

/* Bvh::sample_light + Prim::sample_ray_world. */
static light_sample sample_light(const orc_scene *s, uint64_t seed, uint64_t path, uint32_t stream,
                                 uint32_t bounce, int with_dir)
{
    light_sample ls;
    uint32_t r = (uint32_t)(((uint64_t)orc_draw_u32(seed, path, stream, bounce, ORC_SLOT_LIGHT) * s->num_lights) >> 32);
    uint32_t idx = s->lights[r];
    const orc_prim *p = &s->prims[idx];
    prim_sample_point(p, orc_draw_f32(seed, path, stream, bounce, ORC_SLOT_LIGHT_TRI),
                      orc_draw_f32(seed, path, stream, bounce, ORC_SLOT_LIGHT_U1),
                      orc_draw_f32(seed, path, stream, bounce, ORC_SLOT_LIGHT_U2),
                      &ls.point, &ls.props, &ls.point_pdf);
    ls.point_pdf = ls.point_pdf / (float)s->num_lights;
    ls.prim = idx;
    ls.dir = v_zero();
    ls.dir_pdf = 0.0f;
    if (with_dir) {
        ovec d = orc_cosine_hemisphere(orc_draw_f32(seed, path, stream, bounce, ORC_SLOT_LIGHT_DIR1),
                                       orc_draw_f32(seed, path, stream, bounce, ORC_SLOT_LIGHT_DIR2), 0);
        ls.dir_pdf = v_abs_cos_theta(d) * ORC_FRAC_1_PI;
        ovec t, b;
        v_coord_system(ls.props.geom_normal, &t, &b);
        ls.dir = v_local_to_world(d, t, b, ls.props.geom_normal);
    }
    return ls;
}

/* Area pdf with which sample_light picks a point on prim `pi`. */
static float light_area_pdf(const orc_scene *s, uint32_t pi)
{
    const orc_prim *p = &s->prims[pi];
    float pdf = p->type == ORC_PRIM_MESH ? 1.0f / p->area : 1.0f / (4.0f * ORC_PI * p->radius * p->radius);
    return pdf / (float)s->num_lights;
}

/* power_heuristic(1, f, 1, g), src/core/math.rs:137-142. */
static float power_heuristic(float f_pdf, float g_pdf)
{
    float f = 1.0f * f_pdf, g = 1.0f * g_pdf;
    return (f * f) / (f * f + g * g);
}

/* ---- DisplayColorIntegrator::integrate, integrators.rs:21-34 --------------- */

static ovec display_integrate(const orc_scene *s, ovec o, ovec d)
{
    orc_hit h;
    if (orc_intersect_one(s, o, d, &h, NULL) != 0 || h.prim == ORC_NONE) return v_zero();
    return s->prims[h.prim].mat.display;
}

/* ---- path tracer, integrators.rs:42-99 ------------------------------------ */

static int trace_hit(const orc_scene *s, ovec o, ovec d, orc_hit *h, orc_frame *f)
{
    if (orc_intersect_one(s, o, d, h, NULL) != 0 || h->prim == ORC_NONE) return 0;
    orc_hit_frame(s, h, o, d, f);
    return 1;
}

static void russian_roulette(ovec *T, uint32_t depth, float rv)
{
    float lum = v_luminance(*T);
    float live = depth > 20 ? clamped_lerp(0.10f, 0.75f, lum) : clamped_lerp(0.25f, 1.00f, lum);
    if (rv < live) *T = v_div(*T, live);
    else *T = v_zero();
}

static ovec pt_integrate(const orc_scene *s, ovec o, ovec d, uint64_t seed, uint64_t path,
                         uint32_t max_depth)
{
    uint32_t depth = 0;
    ovec L = v_zero(), T = v_one();
    while (!v_is_exactly_zero(T) && (max_depth == 0 || depth < max_depth)) {
        orc_hit h;
        orc_frame f;
        if (trace_hit(s, o, d, &h, &f)) {
            ovec wi = v_neg(d);
            const orc_material *m = &s->prims[h.prim].mat;
            orc_mat_sample smp = orc_sample_world(
                m, wi, &f, 1, orc_draw_u32(seed, path, ORC_STREAM_CAMERA_PATH, depth, ORC_SLOT_LOBE),
                orc_draw_f32(seed, path, ORC_STREAM_CAMERA_PATH, depth, ORC_SLOT_LOBE_U1),
                orc_draw_f32(seed, path, ORC_STREAM_CAMERA_PATH, depth, ORC_SLOT_LOBE_U2));
            L = v_add(L, v_mul(T, smp.emission));
            T = v_mul(T, v_scale(smp.radiance, fabsf(v_dot(f.normal, smp.outgoing)) / smp.pdf));
            ovec p = v_add(o, v_scale(d, h.dist));
            o = v_add(p, v_scale(smp.outgoing, ORC_RAY_PUSH_DIST));
            d = smp.outgoing;
            if (depth > 10 || v_is_nearly_zero(T))
                russian_roulette(&T, depth, orc_draw_f32(seed, path, ORC_STREAM_CAMERA_PATH, depth, ORC_SLOT_RR));
        } else {
            T = v_zero();
        }
        depth += 1;
    }
    return L;
}

/* ---- path tracer with next-event estimation ------------------------------- */

static ovec nee_integrate(const orc_scene *s, ovec o, ovec d, uint64_t seed, uint64_t path,
                          uint32_t max_depth)
{
    ovec L = v_zero(), T = v_one();
    float prev_pdf = 0.0f;
    int prev_specular = 1;
    for (uint32_t depth = 0; max_depth == 0 || depth < max_depth; ++depth) {
        orc_hit h;
        orc_frame f;
        if (!trace_hit(s, o, d, &h, &f)) break;
        ovec wi = v_neg(d);
        const orc_material *m = &s->prims[h.prim].mat;
        ovec p = v_add(o, v_scale(d, h.dist));
        ovec le = orc_light_world(m, wi, &f);
        if (!v_is_exactly_zero(le)) {
            if (prev_specular) {
                L = v_add(L, v_mul(T, le));
            } else {
                float pl = light_area_pdf(s, h.prim) * (h.dist * h.dist) / fabsf(v_dot(f.geom_normal, wi));
                L = v_add(L, v_scale(v_mul(T, le), power_heuristic(prev_pdf, pl)));
            }
        }
        if (m->nlobes == 0) break;
        if (s->num_lights > 0 && orc_count_lobes(m, LOBE_DIFFUSE | LOBE_GLOSSY) != 0) {
            light_sample ls = sample_light(s, seed, path, ORC_STREAM_CAMERA_PATH, depth, 0);
            ovec to = v_sub(ls.point, p);
            float d2 = v_dot(to, to);
            float dist = sqrtf(d2);
            ovec wo = v_div(to, dist);
            const orc_material *lm = &s->prims[ls.prim].mat;
            ovec le_l = orc_light_world(lm, v_neg(wo), &ls.props);
            float cos_l = fabsf(v_dot(ls.props.geom_normal, wo));
            if (!v_is_exactly_zero(le_l) && cos_l > 0.0f && d2 > 0.0f) {
                ovec fr = orc_f_world(m, wi, wo, &f, 1);
                if (!v_is_exactly_zero(fr)) {
                    float pdf_light = ls.point_pdf * d2 / cos_l;
                    float pdf_bsdf = orc_pdf_world(m, wi, wo, &f);
                    float w = power_heuristic(pdf_light, pdf_bsdf);
                    ovec c = v_scale(v_mul(v_mul(T, fr), le_l), fabsf(v_dot(f.normal, wo)) * w / pdf_light);
                    if (orc_visibility_one(s, p, ls.point, NULL) == 1) L = v_add(L, c);
                }
            }
        }
        orc_mat_sample smp = orc_sample_world(
            m, wi, &f, 1, orc_draw_u32(seed, path, ORC_STREAM_CAMERA_PATH, depth, ORC_SLOT_LOBE),
            orc_draw_f32(seed, path, ORC_STREAM_CAMERA_PATH, depth, ORC_SLOT_LOBE_U1),
            orc_draw_f32(seed, path, ORC_STREAM_CAMERA_PATH, depth, ORC_SLOT_LOBE_U2));
        T = v_mul(T, v_scale(smp.radiance, fabsf(v_dot(f.normal, smp.outgoing)) / smp.pdf));
        prev_pdf = smp.pdf;
        prev_specular = (smp.kind & LOBE_SPECULAR) != 0;
        o = v_add(p, v_scale(smp.outgoing, ORC_RAY_PUSH_DIST));
        d = smp.outgoing;
        if (depth > 10 || v_is_nearly_zero(T))
            russian_roulette(&T, depth, orc_draw_f32(seed, path, ORC_STREAM_CAMERA_PATH, depth, ORC_SLOT_RR));
        if (v_is_exactly_zero(T)) break;
    }
    return L;
}

/* ---- BDPT, integrators.rs:102-418 ---------------------------------------- */

#define BDPT_MAX_DEPTH 16 /* integrators.rs:103 */

typedef struct {
    ovec incoming, point;
    orc_frame props;
    ovec throughput, emission;
    int lobe_kind, connectible;
    uint32_t prim;
    float pdf_forward, pdf_reverse;
    /* MIS (ORC_BDPT_MIS): area-measure densities of reaching this vertex from
     * its predecessor in the walk that made it (a_fwd) and from its successor
     * in the opposite direction (a_rev, set when the successor is sampled). */
    float a_fwd, a_rev;
} bvert; /* BdptVertex, integrators.rs:107-130 */

static float correct_shading_normal(ovec wi, ovec wo, const orc_frame *f, int c2l)
{
    if (c2l) return 1.0f;
    float num = fabsf(v_dot(wi, f->normal)) * fabsf(v_dot(wo, f->geom_normal));
    float den = fabsf(v_dot(wi, f->geom_normal)) * fabsf(v_dot(wo, f->normal));
    return den == 0.0f ? 0.0f : num / den;
}

static float convert_density(ovec prev, ovec next, const orc_frame *nf)
{
    float dist = v_mag(v_sub(next, prev));
    if (dist == 0.0f) return 0.0f;
    ovec dir = v_div(v_sub(next, prev), dist);
    return fabsf(v_dot(nf->normal, dir)) / (dist * dist);
}

/* Solid-angle density at `from` toward `to` as an area density at `to`:
 * |n_to . dir| / dist^2 with the receiving vertex's shading normal. */
static float to_area(ovec from, ovec to, const orc_frame *tf)
{
    return convert_density(from, to, tf);
}

/* BdptIntegrator::random_walk, integrators.rs:138-229.  prev_pdf: the
 * solid-angle density with which the first ray's direction was sampled (the
 * camera's density of its ray, or the light's emission density). */
static void random_walk(const orc_scene *s, ovec o, ovec d, ovec T, int c2l, bvert *st, int *len,
                        int max_len, uint64_t seed, uint64_t path, uint32_t stream, float prev_pdf)
{
    while (!v_is_exactly_zero(T) && *len < max_len) {
        orc_hit h;
        orc_frame f;
        uint32_t b = (uint32_t)*len;
        if (trace_hit(s, o, d, &h, &f)) {
            ovec prev_T = T;
            ovec p = v_add(o, v_scale(d, h.dist));
            ovec wi = v_neg(d);
            const orc_material *m = &s->prims[h.prim].mat;
            orc_mat_sample smp = orc_sample_world(
                m, wi, &f, c2l, orc_draw_u32(seed, path, stream, b, ORC_SLOT_LOBE),
                orc_draw_f32(seed, path, stream, b, ORC_SLOT_LOBE_U1),
                orc_draw_f32(seed, path, stream, b, ORC_SLOT_LOBE_U2));
            float pdf_rev = orc_pdf_world(m, smp.outgoing, wi, &f);
            int conn = orc_count_lobes(m, LOBE_DIFFUSE | LOBE_GLOSSY) != 0;
            T = v_mul(T, v_scale(smp.radiance, fabsf(v_dot(f.normal, smp.outgoing)) / smp.pdf));
            T = v_scale(T, correct_shading_normal(wi, smp.outgoing, &f, c2l));
            float a_fwd;
            if (*len > 0) {
                bvert *pv = &st[*len - 1];
                pv->pdf_reverse = pdf_rev * convert_density(pv->point, p, &f);
                pv->a_rev = pdf_rev * to_area(p, pv->point, &pv->props);
                a_fwd = prev_pdf * to_area(pv->point, p, &f);
            } else {
                a_fwd = prev_pdf * to_area(o, p, &f);  /* the camera's first vertex, from the camera */
            }
            o = v_add(p, v_scale(smp.outgoing, ORC_RAY_PUSH_DIST));
            d = smp.outgoing;
            bvert *v = &st[(*len)++];
            v->incoming = wi; v->point = p; v->props = f; v->throughput = prev_T;
            v->emission = smp.emission; v->lobe_kind = smp.kind; v->connectible = conn;
            v->prim = h.prim; v->pdf_forward = smp.pdf; v->pdf_reverse = 1.0f;
            v->a_fwd = a_fwd; v->a_rev = 0.0f;
            prev_pdf = smp.pdf;
            if (*len >= 4 || v_is_nearly_zero(T)) {
                float rv = orc_draw_f32(seed, path, stream, b, ORC_SLOT_RR);
                float live = clamped_lerp(0.25f, 0.75f, v_luminance(T));
                if (rv < live) T = v_div(T, live);
                else T = v_zero();
            }
        } else {
            T = v_zero();
            bvert *v = &st[(*len)++];
            memset(v, 0, sizeof(*v));
            v->incoming = v_neg(d);
            v->throughput = T;
            v->lobe_kind = LOBE_NONE;
            v->connectible = 0;
            v->prim = ORC_NONE;
            v->pdf_forward = 1.0f; v->pdf_reverse = 1.0f;
            v->a_fwd = 0.0f; v->a_rev = 0.0f;
        }
    }
}

/* BdptIntegrator::connect, integrators.rs:269-343. */
static ovec bdpt_connect(const orc_scene *s, int c, int l, const bvert *cam, const bvert *lig)
{
    const bvert *cv = &cam[c - 1];
    if (l == 0) {
        if (cv->prim == ORC_NONE) return v_zero();
        return v_mul(cv->throughput, cv->emission);
    }
    const bvert *lv = &lig[l - 1];
    if (cv->prim == ORC_NONE || lv->prim == ORC_NONE) return v_zero();
    if (!cv->connectible || !lv->connectible) return v_zero();
    ovec c2l_dir = v_normalized(v_sub(lv->point, cv->point));
    ovec l2c = v_neg(c2l_dir);
    const orc_material *cm = &s->prims[cv->prim].mat, *lm = &s->prims[lv->prim].mat;
    ovec cr = orc_f_world(cm, cv->incoming, c2l_dir, &cv->props, 1);
    ovec ce = l == 1 ? orc_light_world(lm, l2c, &lv->props) : orc_f_world(lm, lv->incoming, l2c, &lv->props, 0);
    float g = fabsf(v_dot(cv->props.normal, c2l_dir));
    float cd = convert_density(cv->point, lv->point, &lv->props);
    ovec contrib = v_scale(v_mul(v_mul(v_mul(cv->throughput, cr), ce), lv->throughput), g * cd);
    if (v_is_nearly_zero(contrib)) return contrib;
    if (orc_visibility_one(s, cv->point, lv->point, NULL) == 1) return contrib;
    return v_zero();
}

/* Density of the light's cosine emission (Prim::sample_ray_world,
 * src/geom/prim.rs:37-51) toward dir from the light point of frame f. */
static float emit_dir_pdf(const orc_frame *f, ovec dir)
{
    float c = v_dot(f->geom_normal, dir);
    return c > 0.0f ? c * ORC_FRAC_1_PI : 0.0f;
}

static float remap0(float p) { return p != 0.0f ? p : 1.0f; }

/* The pinhole camera as the light-tracing strategy sees it: its position
 * (Xform::transform of the origin), the rows of the transform's 3x3 part
 * (camera-space axes in world space) and the window half-extents. */
typedef struct {
    ovec e, ax, ay, az;
    float wx, wy;
    uint32_t width, height;
} lt_camera;

/* Solid-angle density of one camera ray through a window point whose
 * camera-space direction makes cos_theta with the view axis: sample points
 * cover the film with density 1 per pixel area and iteration (a pixel's
 * point is uniform over the 4x4 pixels around it), i.e. 1 / (W H) per path
 * over the window of area (2 wx)(2 wy) at distance 1, whose solid angle
 * element is cos^3 theta dA. */
static float camera_dir_pdf(float wx, float wy, float cos_theta)
{
    return 1.0f / (((4.0f * wx) * wy) * ((cos_theta * cos_theta) * cos_theta));
}

/* Projection of a world point seen from the camera: the film position
 * (col_cont - 0.5, row_cont - 0.5) as report_samples computes it for a
 * sample, the unit direction from the point to the camera, the distance and
 * cos theta.  0 when the point is behind the camera or its splat (the pixels
 * within the filter width) misses the film. */
static int lt_project(const lt_camera *c, ovec y, float *cd, float *rd, ovec *to_cam, float *dist, float *cos_t)
{
    const ovec d = v_sub(y, c->e);
    const float len = v_mag(d);
    if (!(len > 0.0f)) return 0;
    const ovec u = v_div(d, len);
    const float dx = v_dot(u, c->ax), dy = v_dot(u, c->ay), dz = v_dot(u, c->az);
    if (!(dz < 0.0f)) return 0;
    const float s = (dx / -dz) / c->wx, t = (dy / -dz) / c->wy;
    const float wf = (float)c->width, hf = (float)c->height;
    const float col_cont = lerpf(0.0f, wf, 0.5f * (s + 1.0f));
    const float row_cont = lerpf(0.0f, hf, 0.5f * (t + 1.0f));
    *cd = col_cont - 0.5f;
    *rd = row_cont - 0.5f;
    if (!(floorf(*cd + FILTER_WIDTH) >= 0.0f && ceilf(*cd - FILTER_WIDTH) <= wf - 1.0f &&
          floorf(*rd + FILTER_WIDTH) >= 0.0f && ceilf(*rd - FILTER_WIDTH) <= hf - 1.0f))
        return 0;
    *to_cam = v_neg(u);
    *dist = len;
    *cos_t = -dz / sqrtf(dx * dx + dy * dy + dz * dz);
    return 1;
}

/* Power-heuristic MIS weight of strategy (c, l) among the strategies that
 * can make the same path z_1..z_k (k = c + l; z_i = camera vertex i for
 * i <= c, z_{c+m} = light vertex l-m+1): j camera vertices and k-j light
 * vertices, j = 1..k, within the depth caps, connecting two connectible
 * vertices (j < k) or hitting the emitter (j = k), with the light point z_k
 * on a light.  p_j = prod_{i<=j} pF(z_i) prod_{i>j} pL(z_i); the weight is
 * p_c^2 / sum_j p_j^2 from the ratios p_{j-1}/p_j = pL(z_j)/pF(z_j).
 * With a light-tracing camera (ltc != NULL) j = 0 is a strategy too (all k
 * vertices from the light walk, z_1 joined to the camera, which counts as
 * connectible; pF(z_1) is the camera's density), and c = 0 is allowed:
 * z_1 = y_l seen directly by the camera. */
static float bdpt_mis_weight(const orc_scene *s, int c, int l, const bvert *cam, const bvert *lig, int cmax,
                             int lmax, const lt_camera *ltc)
{
    const int k = c + l;
    float pf[2 * BDPT_MAX_DEPTH + 2], pl[2 * BDPT_MAX_DEPTH + 2];
    int conn[2 * BDPT_MAX_DEPTH + 2];
    conn[0] = 1;
    for (int i = 1; i <= c; ++i) {
        pf[i] = cam[i - 1].a_fwd;
        pl[i] = cam[i - 1].a_rev;
        conn[i] = cam[i - 1].connectible;
    }
    for (int m = 1; m <= l; ++m) {
        const bvert *y = &lig[l - m];  /* light vertex l-m+1 */
        pf[c + m] = y->a_rev;
        pl[c + m] = y->a_fwd;
        conn[c + m] = y->connectible;
    }
    int emissive;
    if (c == 0) {  /* light tracing: y_l joined to the camera */
        emissive = 1;
        const bvert *yl = &lig[l - 1];
        float cd, rd, dist, cos_t;
        ovec d_yc;
        if (!lt_project(ltc, yl->point, &cd, &rd, &d_yc, &dist, &cos_t)) return 0.0f;
        pf[1] = camera_dir_pdf(ltc->wx, ltc->wy, cos_t) * to_area(ltc->e, yl->point, &yl->props);
        if (l >= 2) {
            const orc_material *my = &s->prims[yl->prim].mat;
            const bvert *yp = &lig[l - 2];
            pf[2] = orc_pdf_world(my, d_yc, yl->incoming, &yl->props) * to_area(yl->point, yp->point, &yp->props);
        }
    } else if (l == 0) {
        const bvert *xc = &cam[c - 1];
        emissive = xc->prim != ORC_NONE && s->prims[xc->prim].mat.has_light;
        if (!emissive) return 0.0f;
        pl[k] = light_area_pdf(s, xc->prim);
        conn[k] = 1;  /* as a light point it can be connected to */
        if (c >= 2) {
            const bvert *xp = &cam[c - 2];
            pl[c - 1] = emit_dir_pdf(&xc->props, v_normalized(v_sub(xp->point, xc->point))) *
                        to_area(xc->point, xp->point, &xp->props);
        }
    } else {
        emissive = 1;
        const bvert *xc = &cam[c - 1];
        const bvert *yl = &lig[l - 1];
        const orc_material *mc = &s->prims[xc->prim].mat;
        const ovec d_cy = v_normalized(v_sub(yl->point, xc->point)), d_yc = v_neg(d_cy);
        pf[c + 1] = orc_pdf_world(mc, xc->incoming, d_cy, &xc->props) * to_area(xc->point, yl->point, &yl->props);
        if (l == 1) {
            pl[c] = emit_dir_pdf(&yl->props, d_yc) * to_area(yl->point, xc->point, &xc->props);
        } else {
            const orc_material *my = &s->prims[yl->prim].mat;
            pl[c] = orc_pdf_world(my, yl->incoming, d_yc, &yl->props) * to_area(yl->point, xc->point, &xc->props);
            const bvert *yp = &lig[l - 2];
            pf[c + 2] = orc_pdf_world(my, d_yc, yl->incoming, &yl->props) * to_area(yl->point, yp->point, &yp->props);
        }
        if (c >= 2) {
            const bvert *xp = &cam[c - 2];
            pl[c - 1] = orc_pdf_world(mc, d_cy, xc->incoming, &xc->props) * to_area(xc->point, xp->point, &xp->props);
        }
    }
    float sum = 0.0f, r = 1.0f;
    const int jmin = ltc ? 0 : 1;
#define ORC_MIS_VALID(j) ((j) <= cmax && k - (j) <= lmax && emissive && ((j) == k || (conn[j] && conn[(j) + 1])))
    if (ORC_MIS_VALID(c)) sum += 1.0f;
    for (int j = c - 1; j >= jmin; --j) {
        r = r * (remap0(pl[j + 1]) / remap0(pf[j + 1]));
        if (ORC_MIS_VALID(j)) sum += r * r;
    }
    r = 1.0f;
    for (int j = c + 1; j <= k; ++j) {
        r = r * (remap0(pf[j]) / remap0(pl[j]));
        if (ORC_MIS_VALID(j)) sum += r * r;
    }
#undef ORC_MIS_VALID
    return sum > 0.0f ? 1.0f / sum : 0.0f;
}

/* One light-tracing contribution: a light vertex's radiance toward the
 * camera, to be splatted at film position (cd, rd). */
typedef struct {
    float cd, rd;
    ovec color;
} lt_record;

/* The light-tracing strategy (c = 0) for light vertex y: radiance carried by
 * the light subpath to y and scattered toward the camera, T f |n.w| / d^2,
 * times the camera's density of that direction (the importance of a pinhole
 * whose film is sampled uniformly), occluded or not, times its MIS weight. */
static int light_trace(const orc_scene *s, int l, const bvert *cam, const bvert *lig, int cmax, int lmax,
                       const lt_camera *ltc, lt_record *out)
{
    const bvert *y = &lig[l - 1];
    if (y->prim == ORC_NONE || !y->connectible) return 0;
    float cd, rd, dist, cos_t;
    ovec to_cam;
    if (!lt_project(ltc, y->point, &cd, &rd, &to_cam, &dist, &cos_t)) return 0;
    const orc_material *m = &s->prims[y->prim].mat;
    ovec f = l == 1 ? orc_light_world(m, to_cam, &y->props) : orc_f_world(m, y->incoming, to_cam, &y->props, 0);
    float g = convert_density(ltc->e, y->point, &y->props);
    ovec contrib = v_scale(v_mul(y->throughput, f), g * camera_dir_pdf(ltc->wx, ltc->wy, cos_t));
    if (v_is_nearly_zero(contrib)) return 0;
    if (orc_visibility_one(s, ltc->e, y->point, NULL) != 1) return 0;
    float w = bdpt_mis_weight(s, 0, l, cam, lig, cmax, lmax, ltc);
    out->cd = cd;
    out->rd = rd;
    out->color = v_scale(contrib, w);
    return 1;
}

/* lt: the light-tracing camera (NULL: camera strategies only); lt_out gets
 * up to lmax light-tracing records, *lt_n their count. */
static ovec bdpt_integrate(const orc_scene *s, ovec o, ovec d, float cam_pdf, uint64_t seed, uint64_t path,
                           uint32_t max_depth, uint32_t light_depth, int mis, const lt_camera *lt,
                           lt_record *lt_out, int *lt_n)
{
    bvert cam[BDPT_MAX_DEPTH], lig[BDPT_MAX_DEPTH];
    int nc = 0, nl = 0;
    int cmax = (max_depth == 0 || max_depth > BDPT_MAX_DEPTH) ? BDPT_MAX_DEPTH : (int)max_depth;
    int lmax = (light_depth == 0 || light_depth + 1 > BDPT_MAX_DEPTH) ? BDPT_MAX_DEPTH : (int)light_depth + 1;
    ovec L = v_zero();
    if (lt_n) *lt_n = 0;
    random_walk(s, o, d, v_one(), 1, cam, &nc, cmax, seed, path, ORC_STREAM_CAMERA_PATH, cam_pdf);
    if (nc == 0 || s->num_lights == 0) return L;
    light_sample ls = sample_light(s, seed, path, ORC_STREAM_LIGHT_PATH, 0, 1);
    if (ls.point_pdf == 0.0f || ls.dir_pdf == 0.0f) return L;
    const orc_material *lm = &s->prims[ls.prim].mat;
    ovec init = v_scale(orc_light_world(lm, ls.dir, &ls.props),
                        fabsf(v_dot(ls.props.geom_normal, ls.dir)) / (ls.point_pdf * ls.dir_pdf));
    bvert *v0 = &lig[nl++];
    memset(v0, 0, sizeof(*v0));
    v0->incoming = v_zero(); v0->point = ls.point; v0->props = ls.props;
    v0->throughput = v_div(v_one(), ls.point_pdf);
    v0->emission = v_zero(); v0->lobe_kind = LOBE_NONE; v0->connectible = 1; v0->prim = ls.prim;
    v0->pdf_forward = ls.point_pdf * ls.dir_pdf; v0->pdf_reverse = 1.0f;
    v0->a_fwd = ls.point_pdf; v0->a_rev = 0.0f;
    random_walk(s, v_add(ls.point, v_scale(ls.dir, ORC_RAY_PUSH_DIST)), ls.dir, init, 0, lig, &nl, lmax,
                seed, path, ORC_STREAM_LIGHT_PATH, ls.dir_pdf);
    for (int c = 1; c < nc + 1; ++c) {
        for (int l = 0; l < nl + 1; ++l) {
            if (c + l > BDPT_MAX_DEPTH) continue;
            ovec con = bdpt_connect(s, c, l, cam, lig);
            if (mis && v_is_exactly_zero(con)) continue;
            float w = mis ? bdpt_mis_weight(s, c, l, cam, lig, cmax, lmax, lt)
                          : 1.0f / (float)(c + l); /* weight, integrators.rs:345-354 */
            L = v_add(L, v_scale(con, w));
        }
    }
    if (lt) {
        for (int l = 1; l <= nl; ++l)
            if (light_trace(s, l, cam, lig, cmax, lmax, lt, &lt_out[*lt_n])) ++*lt_n;
    }
    return L;
}

/* Adds a light-tracing record to the per-pixel splat sums with the film's
 * Mitchell weights over the pixels within the filter width (the footprint
 * report_samples gives a sample at the same position).  Each pixel's sum
 * estimates the filter-weighted radiance over the film per iteration. */
static void lt_splat(float *splat, uint32_t w, uint32_t h, const lt_record *r)
{
    int64_t last_col = (int64_t)w - 1, last_row = (int64_t)h - 1;
    int64_t min_col = clamp_i(f2i(ceilf(r->cd - FILTER_WIDTH)), 0, last_col);
    int64_t max_col = clamp_i(f2i(floorf(r->cd + FILTER_WIDTH)), 0, last_col);
    int64_t min_row = clamp_i(f2i(ceilf(r->rd - FILTER_WIDTH)), 0, last_row);
    int64_t max_row = clamp_i(f2i(floorf(r->rd + FILTER_WIDTH)), 0, last_row);
    for (int64_t y = min_row; y < max_row + 1; ++y) {
        for (int64_t x = min_col; x < max_col + 1; ++x) {
            float *px = &splat[3 * (y * (int64_t)w + x)];
            float wt = mitchell2((float)x - r->cd, (float)y - r->rd);
            px[0] = px[0] + r->color.x * wt;
            px[1] = px[1] + r->color.y * wt;
            px[2] = px[2] + r->color.z * wt;
        }
    }
}

/* Antiderivative of the 1-D Mitchell kernel k(t) = mitchell_filter1(t / 2)
 * (B = C = 1/3, support |t| <= 2, integral 1), odd, in double. */
static double mitchell_integral(double t)
{
    const double a = fabs(t);
    double r;
    if (a <= 1.0) {
        r = (1.75 * a * a * a * a - 4.0 * a * a * a + (16.0 / 3.0) * a) / 6.0;
    } else {
        const double b = a < 2.0 ? a : 2.0;
#define ORC_MITCHELL_F(u) (-(7.0 / 12.0) * (u) * (u) * (u) * (u) + 4.0 * (u) * (u) * (u) - 10.0 * (u) * (u) + (32.0 / 3.0) * (u))
        r = (1.75 - 4.0 + 16.0 / 3.0) / 6.0 + (ORC_MITCHELL_F(b) - ORC_MITCHELL_F(1.0)) / 6.0;
#undef ORC_MITCHELL_F
    }
    return t < 0.0 ? -r : r;
}

/* Expected film weight one iteration's samples give pixel column x of n
 * (rows alike): the sample of pixel q lies uniformly within q +- 2, so
 * E = (1/4) sum_q int_{-2}^{2} k(x - q - j) dj -- 1 inside the film, less
 * within two pixels of its edges.  A pixel's expectation is the product of
 * its column's and its row's. */
void orc_expected_weights(uint32_t n, float *out)
{
    for (int64_t x = 0; x < (int64_t)n; ++x) {
        double e = 0.0;
        for (int64_t q = x - 4; q <= x + 4; ++q) {
            if (q < 0 || q >= (int64_t)n) continue;
            e += mitchell_integral((double)(x - q) + 2.0) - mitchell_integral((double)(x - q) - 2.0);
        }
        out[x] = (float)(0.25 * e);
    }
}

/* ---- render driver (Stage::trace, src/render/stage.rs:23-36) -------------- */

int orc_render(const orc_scene *s, const float cam[16], float wx, float wy, uint32_t width,
               uint32_t height, int integrator, uint32_t max_depth, uint32_t light_depth,
               uint64_t seed, uint64_t sample_offset, uint32_t spp, float *film, float *colors,
               int nthreads)
{
    if (!s->nodes || width == 0 || height == 0) return -1;
    uint64_t npx = (uint64_t)width * height;
    float *st = (float *)malloc(sizeof(float) * 2 * npx);
    float *col = (float *)malloc(sizeof(float) * 3 * npx);
    if (!st || !col) { free(st); free(col); return -1; }
    if (nthreads <= 0) nthreads = 1;
    /* light tracing (ORC_BDPT_LT): per path up to lmax splat records, summed
     * per pixel in path order after each iteration */
    const int lt_on = integrator == ORC_BDPT_LT;
    const int lmax = (light_depth == 0 || light_depth + 1 > BDPT_MAX_DEPTH) ? BDPT_MAX_DEPTH : (int)light_depth + 1;
    lt_camera ltc;
    lt_record *ltr = NULL;
    int *ltn = NULL;
    float *splat = NULL, *ew_col = NULL, *ew_row = NULL;
    if (lt_on) {
        ltc.e = v_new(cam[12] / cam[15], cam[13] / cam[15], cam[14] / cam[15]);
        ltc.ax = v_new(cam[0], cam[1], cam[2]);
        ltc.ay = v_new(cam[4], cam[5], cam[6]);
        ltc.az = v_new(cam[8], cam[9], cam[10]);
        ltc.wx = wx; ltc.wy = wy; ltc.width = width; ltc.height = height;
        ltr = (lt_record *)malloc(sizeof(lt_record) * npx * (size_t)lmax);
        ltn = (int *)malloc(sizeof(int) * npx);
        splat = (float *)malloc(sizeof(float) * 3 * npx);
        ew_col = (float *)malloc(sizeof(float) * width);
        ew_row = (float *)malloc(sizeof(float) * height);
        if (!ltr || !ltn || !splat || !ew_col || !ew_row) {
            free(st); free(col); free(ltr); free(ltn); free(splat); free(ew_col); free(ew_row);
            return -1;
        }
        orc_expected_weights(width, ew_col);
        orc_expected_weights(height, ew_row);
    }
    for (uint32_t it = 0; it < spp; ++it) {
        uint64_t iter = sample_offset + it;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)npx; ++i) {
            uint64_t path = iter * npx + (uint64_t)i;
            uint32_t row = (uint32_t)(i / width), colm = (uint32_t)(i % width);
            float ss, tt;
            sample_point(seed, path, row, colm, width, height, &ss, &tt);
            st[2 * i] = ss; st[2 * i + 1] = tt;
            ovec o, d;
            camera_ray(cam, wx, wy, ss, tt, &o, &d);
            const ovec cdir = v_normalized(v_new(wx * ss, wy * tt, -1.0f));
            const float cam_pdf = camera_dir_pdf(wx, wy, -cdir.z);
            ovec L;
            if (integrator == 0) L = pt_integrate(s, o, d, seed, path, max_depth);
            else if (integrator == 1) L = nee_integrate(s, o, d, seed, path, max_depth);
            else if (integrator == ORC_DISPLAY) L = display_integrate(s, o, d);
            else if (integrator == ORC_BDPT_MIS || integrator == ORC_BDPT_LT)
                L = bdpt_integrate(s, o, d, cam_pdf, seed, path, max_depth, light_depth, 1, lt_on ? &ltc : NULL,
                                   lt_on ? &ltr[(size_t)i * lmax] : NULL, lt_on ? &ltn[i] : NULL);
            else L = bdpt_integrate(s, o, d, cam_pdf, seed, path, max_depth, light_depth, 0, NULL, NULL, NULL);
            col[3 * i] = L.x; col[3 * i + 1] = L.y; col[3 * i + 2] = L.z;
        }
        if (colors) memcpy(&colors[3 * npx * it], col, sizeof(float) * 3 * npx);
        if (lt_on) {
            memset(splat, 0, sizeof(float) * 3 * npx);
            for (uint64_t i = 0; i < npx; ++i)
                for (int r = 0; r < ltn[i]; ++r) lt_splat(splat, width, height, &ltr[i * lmax + r]);
        }
        if (film) {
            report_samples(film, width, height, st, col, npx);
            if (lt_on) {
                /* this iteration's light-traced estimate of the pixel's filtered
                 * radiance, weighted by the film weight the iteration's camera
                 * samples give the pixel in expectation */
                for (uint64_t p = 0; p < npx; ++p) {
                    const float ew = ew_col[p % width] * ew_row[p / width];
                    film[4 * p] = film[4 * p] + ew * splat[3 * p];
                    film[4 * p + 1] = film[4 * p + 1] + ew * splat[3 * p + 1];
                    film[4 * p + 2] = film[4 * p + 2] + ew * splat[3 * p + 2];
                }
            }
        }
    }
    free(st);
    free(col);
    free(ltr); free(ltn); free(splat); free(ew_col); free(ew_row);
    return 0;
}

void orc_sample_points(uint64_t seed, uint64_t sample_offset, uint32_t width, uint32_t height, float *st)
{
    uint64_t npx = (uint64_t)width * height;
    for (uint64_t i = 0; i < npx; ++i)
        sample_point(seed, sample_offset * npx + i, (uint32_t)(i / width), (uint32_t)(i % width), width,
                     height, &st[2 * i], &st[2 * i + 1]);
}

void orc_report_samples(float *film, uint32_t width, uint32_t height, const float *st, const float *colors,
                        int64_t n)
{
    report_samples(film, width, height, st, colors, (uint64_t)n);
}

/* ---- EXR, src/render/exr.rs ------------------------------------------------ */

typedef struct { uint8_t *p; size_t n, cap; } obuf;
static void ob_raw(obuf *b, const void *src, size_t n)
{
    if (b->p && b->n + n <= b->cap) memcpy(b->p + b->n, src, n);
    b->n += n;
}
static void ob_i32(obuf *b, int32_t v) { ob_raw(b, &v, 4); }
static void ob_u64(obuf *b, uint64_t v) { ob_raw(b, &v, 8); }
static void ob_f32(obuf *b, float v) { ob_raw(b, &v, 4); }
static void ob_u8(obuf *b, uint8_t v) { ob_raw(b, &v, 1); }
static void ob_str(obuf *b, const char *s) { ob_raw(b, s, strlen(s) + 1); }

/* ExrWriter::update (exr.rs:168-194): header, line offset table, then
 * write_channels (exr.rs:138-166).  Writes up to cap bytes into out (may be
 * NULL) and returns the full size. */
size_t orc_exr_encode(const float *rgbw, uint32_t width, uint32_t height, uint8_t *out, size_t cap)
{
    obuf b = {out, 0, cap};
    ob_i32(&b, 20000630);
    ob_i32(&b, 2);
    ob_str(&b, "channels"); ob_str(&b, "chlist");
    ob_i32(&b, 2 * 3 + 16 * 3 + 1);
    const char *ch[3] = {"B", "G", "R"};
    for (int k = 0; k < 3; ++k) {
        ob_str(&b, ch[k]); ob_i32(&b, 2); ob_i32(&b, 0); ob_i32(&b, 1); ob_i32(&b, 1);
    }
    ob_u8(&b, 0);
    ob_str(&b, "compression"); ob_str(&b, "compression"); ob_i32(&b, 1); ob_u8(&b, 0);
    int32_t win[4] = {0, 0, (int32_t)width - 1, (int32_t)height - 1};
    ob_str(&b, "dataWindow"); ob_str(&b, "box2i"); ob_i32(&b, 16);
    for (int k = 0; k < 4; ++k) ob_i32(&b, win[k]);
    ob_str(&b, "displayWindow"); ob_str(&b, "box2i"); ob_i32(&b, 16);
    for (int k = 0; k < 4; ++k) ob_i32(&b, win[k]);
    ob_str(&b, "lineOrder"); ob_str(&b, "lineOrder"); ob_i32(&b, 1); ob_u8(&b, 0);
    ob_str(&b, "pixelAspectRatio"); ob_str(&b, "float"); ob_i32(&b, 4); ob_f32(&b, 1.0f);
    ob_str(&b, "screenWindowCenter"); ob_str(&b, "v2f"); ob_i32(&b, 8); ob_f32(&b, 0.0f); ob_f32(&b, 0.0f);
    ob_str(&b, "screenWindowWidth"); ob_str(&b, "float"); ob_i32(&b, 4); ob_f32(&b, (float)width);
    ob_u8(&b, 0);
    size_t line_size = 4 + 4 + (size_t)width * 4 * 3;
    size_t data_offset = b.n + 8 * (size_t)height;
    for (uint32_t y = 0; y < height; ++y) ob_u64(&b, (uint64_t)(data_offset + y * line_size));
    for (uint32_t y = 0; y < height; ++y) {
        ob_i32(&b, (int32_t)y);
        uint32_t ls = (uint32_t)line_size - 8;
        ob_raw(&b, &ls, 4);
        const float *row = &rgbw[4 * (size_t)(height - y - 1) * width];
        for (uint32_t i = 0; i < width; ++i) ob_f32(&b, row[4 * i + 2] / row[4 * i + 3]);
        for (uint32_t i = 0; i < width; ++i) ob_f32(&b, row[4 * i + 1] / row[4 * i + 3]);
        for (uint32_t i = 0; i < width; ++i) ob_f32(&b, row[4 * i + 0] / row[4 * i + 3]);
    }
    return b.n;
}

/* Per-function material evaluation for tests.  in: 24 floats per item --
 * wi(3), wo(3), frame normal/tangent/binormal/geom_normal (12), u_lobe (u32
 * bits), u1, u2, camera_to_light, pad(2).  out: 20 floats per item --
 * f_world(3), pdf_world, light_world(3), sample emission(3), radiance(3),
 * outgoing(3), pdf, kind, count_lobes(DIFFUSE|GLOSSY), density of the draws at wo. */
void orc_material_query(const orc_scene *s, uint32_t prim, int64_t n, const float *in, float *out)
{
    const orc_material *m = &s->prims[prim].mat;
    for (int64_t i = 0; i < n; ++i) {
        const float *q = &in[24 * i];
        float *r = &out[20 * i];
        ovec wi = v_new(q[0], q[1], q[2]), wo = v_new(q[3], q[4], q[5]);
        orc_frame f;
        f.normal = v_new(q[6], q[7], q[8]);
        f.tangent = v_new(q[9], q[10], q[11]);
        f.binormal = v_new(q[12], q[13], q[14]);
        f.geom_normal = v_new(q[15], q[16], q[17]);
        uint32_t ul;
        memcpy(&ul, &q[18], 4);
        int c2l = q[21] != 0.0f;
        ovec fw = orc_f_world(m, wi, wo, &f, c2l);
        float pw = orc_pdf_world(m, wi, wo, &f);
        ovec lw = orc_light_world(m, wi, &f);
        orc_mat_sample sm = orc_sample_world(m, wi, &f, c2l, ul, q[19], q[20]);
        r[0] = fw.x; r[1] = fw.y; r[2] = fw.z; r[3] = pw;
        r[4] = lw.x; r[5] = lw.y; r[6] = lw.z;
        r[7] = sm.emission.x; r[8] = sm.emission.y; r[9] = sm.emission.z;
        r[10] = sm.radiance.x; r[11] = sm.radiance.y; r[12] = sm.radiance.z;
        r[13] = sm.outgoing.x; r[14] = sm.outgoing.y; r[15] = sm.outgoing.z;
        r[16] = sm.pdf; r[17] = (float)sm.kind;
        r[18] = (float)orc_count_lobes(m, LOBE_DIFFUSE | LOBE_GLOSSY);
        r[19] = orc_sample_density_world(m, wi, wo, &f);
    }
}
