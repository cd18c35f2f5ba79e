/*
 * orc_lobes.c -- CPU oracle: samplers, Fresnel and microfacet models, Disney
 * lobes and the material dispatch (test infrastructure only; see gr_oracle.h).
 */
#include "orc_render.h"

#include <string.h>

/* ---- samplers, src/core/random.rs ---------------------------------------- */

/* AreaSampleDisk with its two Range(-1, 1) draws given as u1, u2 in [0, 1)
 * (src/core/random.rs:37-58). */
void orc_area_sample_disk(float u1, float u2, float *x, float *y)
{
    float sx = -1.0f + 2.0f * u1;
    float sy = -1.0f + 2.0f * u2;
    if (sx == 0.0f && sy == 0.0f) { *x = 0.0f; *y = 0.0f; return; }
    float r, theta;
    if (fabsf(sx) > fabsf(sy)) {
        r = sx;
        theta = ORC_FRAC_PI_4 * (sy / sx);
    } else {
        r = sy;
        theta = ORC_FRAC_PI_2 - ORC_FRAC_PI_4 * (sx / sy);
    }
    *x = r * cosf(theta);
    *y = r * sinf(theta);
}

/* CosineSampleHemisphere, src/core/random.rs:93-106. */
ovec orc_cosine_hemisphere(float u1, float u2, int flipped)
{
    float x, y;
    orc_area_sample_disk(u1, u2, &x, &y);
    float z = sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
    return flipped ? v_new(x, y, -1.0f * z) : v_new(x, y, z);
}

/* Uniform point on the unit sphere.  The reference normalises three
 * StandardNormal draws (src/core/random.rs:131-150); this draws the same
 * distribution by inversion: z = 1 - 2 u1, phi = 2 pi u2. */
ovec orc_uniform_sphere(float u1, float u2)
{
    float z = 1.0f - 2.0f * u1;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float phi = ORC_TWO_PI * u2;
    return v_new(r * cosf(phi), r * sinf(phi), z);
}

/* ---- Fresnel, src/material/util.rs:8-110 --------------------------------- */

static float schlick_weight(float cos_theta)
{
    float x = clamp_unit(1.0f - cos_theta);
    return x * x * x * x * x;
}

static ovec fresnel_schlick(float cos_theta, ovec r0)
{
    return v_lerp(r0, v_one(), schlick_weight(cos_theta));
}

static float fresnel_dielectric(float cos_in, float ior)
{
    int entering = cos_in > 0.0f;
    float eta_i, eta_t, c;
    if (entering) { eta_i = 1.0f; eta_t = ior; c = clamp_unit(cos_in); }
    else { eta_i = ior; eta_t = 1.0f; c = clamp_unit(-cos_in); }
    float sin_in = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float sin_tr = eta_i / eta_t * sin_in;
    if (sin_tr >= 1.0f) return 1.0f;
    float cos_tr = sqrtf(fmaxf(0.0f, 1.0f - sin_tr * sin_tr));
    float r_parl = ((eta_t * c) - (eta_i * cos_tr)) / ((eta_t * c) + (eta_i * cos_tr));
    float r_perp = ((eta_i * c) - (eta_t * cos_tr)) / ((eta_i * c) + (eta_t * cos_tr));
    return (r_parl * r_parl + r_perp * r_perp) / 2.0f;
}

/* DisneyFresnel::fresnel, util.rs:72-78. */
static ovec disney_fresnel(const orc_lobe *l, float cos_theta)
{
    ovec dielectric = v_scale(l->fr_spec_color, fresnel_dielectric(cos_theta, l->fr_ior));
    ovec conductor = fresnel_schlick(cos_theta, l->fr_color);
    return v_lerp(dielectric, conductor, l->fr_metallic);
}

/* ---- GGX, util.rs:119-251 ------------------------------------------------- */

static float ggx_lambda(const orc_lobe *l, ovec v)
{
    float abs_tan = fabsf(v_tan_theta(v));
    if (isfinite(abs_tan)) {
        float alpha = sqrtf(v_cos2_phi(v) * (l->ax * l->ax) + v_sin2_phi(v) * (l->ay * l->ay));
        float a2t2 = (alpha * abs_tan) * (alpha * abs_tan);
        return (-1.0f + sqrtf(1.0f + a2t2)) * 0.5f;
    }
    return 0.0f;
}

static float ggx_g1(const orc_lobe *l, ovec v) { return 1.0f / (1.0f + ggx_lambda(l, v)); }

static float ggx_d(const orc_lobe *l, ovec h)
{
    float tan2 = v_tan2_theta(h);
    if (isfinite(tan2)) {
        float cos4 = v_cos2_theta(h) * v_cos2_theta(h);
        float e = (v_cos2_phi(h) / (l->ax * l->ax) + v_sin2_phi(h) / (l->ay * l->ay)) * tan2;
        return 1.0f / (ORC_PI * l->ax * l->ay * cos4 * (1.0f + e) * (1.0f + e));
    }
    return 0.0f;
}

static float ggx_g(const orc_lobe *l, ovec i, ovec o)
{
    return 1.0f / (1.0f + ggx_lambda(l, i) + ggx_lambda(l, o));
}

/* GgxDistribution::sample11, util.rs:152-194. */
static void ggx_sample11(float cos_theta, float u1, float u2, float *sx, float *sy)
{
    if (cos_theta > 0.9999f) {
        float r = sqrtf(u1 / (1.0f - u1));
        float phi = ORC_TWO_PI * u2;
        *sx = r * cosf(phi);
        *sy = r * sinf(phi);
        return;
    }
    float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
    float tan_theta = sin_theta / cos_theta;
    float g1 = 2.0f / (1.0f + sqrtf(1.0f + tan_theta * tan_theta));
    float a = 2.0f * u1 / g1 - 1.0f;
    float tmp = fminf(1.0f / (a * a - 1.0f), 1e10f);
    float b = tan_theta;
    float d = sqrtf(fmaxf(b * b * tmp * tmp - (a * a - b * b) * tmp, 0.0f));
    float slope_x_1 = b * tmp - d;
    float slope_x_2 = b * tmp + d;
    float slope_x = (a < 0.0f || slope_x_2 > 1.0f / tan_theta) ? slope_x_1 : slope_x_2;
    float s, u;
    if (u2 > 0.5f) { s = 1.0f; u = 2.0f * (u2 - 0.5f); }
    else { s = -1.0f; u = 2.0f * (0.5f - u2); }
    float z = (u * (u * (u * 0.27385f - 0.73369f) + 0.46341f)) /
              (u * (u * (u * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
    *sx = slope_x;
    *sy = s * z * sqrtf(1.0f + slope_x * slope_x);
}

/* GgxDistribution::sample_half, util.rs:214-240. */
static ovec ggx_sample_half(const orc_lobe *l, ovec i, float u1, float u2)
{
    int flip = i.z < 0.0f;
    ovec f = flip ? v_neg(i) : i;
    ovec st = v_normalized(v_new(l->ax * f.x, l->ay * f.y, f.z));
    float cos_theta = st.z;
    float sx, sy;
    ggx_sample11(cos_theta, u1, u2, &sx, &sy);
    float cos_phi = v_cos_phi(st), sin_phi = v_sin_phi(st);
    float rx = l->ax * (cos_phi * sx - sin_phi * sy);
    float ry = l->ay * (sin_phi * sx + cos_phi * sy);
    ovec h = v_normalized(v_new(-rx, -ry, 1.0f));
    return flip ? v_neg(h) : h;
}

/* GgxDistribution::pdf, util.rs:242-250. */
static float ggx_pdf(const orc_lobe *l, ovec i, ovec h)
{
    float c = i.z;
    if (c == 0.0f) return 0.0f;
    return ggx_d(l, h) * ggx_g1(l, i) * fabsf(v_dot(i, h)) / fabsf(c);
}

/* ---- GTR1, util.rs:255-309 (Q21) ------------------------------------------ */

static float gtr1_lambda(ovec v)
{
    float alpha_g = 0.25f;
    float c = fabsf(v.z);
    float a2 = alpha_g * alpha_g;
    float c2 = c * c;
    return 1.0f / (c + sqrtf(a2 + c2 - (a2 * c2)));
}

static float gtr1_d(const orc_lobe *l, ovec h)
{
    float a2 = l->alpha * l->alpha;
    float c = fabsf(h.z);
    return (a2 - 1.0f) / (ORC_PI * logf(a2) * (1.0f + (a2 - 1.0f) * c * c));
}

static float gtr1_g(ovec i, ovec o) { return 1.0f / (1.0f + gtr1_lambda(i) + gtr1_lambda(o)); }

static ovec gtr1_sample_half(const orc_lobe *l, ovec i, float u1, float u2)
{
    float a2 = l->alpha * l->alpha;
    float phi = 2.0f * ORC_PI * u1;
    float cos_theta = sqrtf(clamp_unit(1.0f - powf(a2, 1.0f - u2) / (1.0f - a2)));
    /* Vec::from_spherical, vector.rs:175-178 */
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    ovec h = v_new(sin_theta * cosf(phi), sin_theta * sinf(phi), cos_theta);
    return v_same_hemisphere(h, i) ? h : v_neg(h);
}

/* ---- microfacet distribution / Fresnel dispatch --------------------------- */

static float dist_d(const orc_lobe *l, ovec h)
{
    return l->type == OL_CLEARCOAT ? gtr1_d(l, h) : ggx_d(l, h);
}
static float dist_g(const orc_lobe *l, ovec i, ovec o)
{
    return l->type == OL_CLEARCOAT ? gtr1_g(i, o) : ggx_g(l, i, o);
}
static float dist_pdf(const orc_lobe *l, ovec i, ovec h)
{
    return l->type == OL_CLEARCOAT ? gtr1_d(l, h) : ggx_pdf(l, i, h);
}
static ovec dist_sample_half(const orc_lobe *l, ovec i, float u1, float u2)
{
    return l->type == OL_CLEARCOAT ? gtr1_sample_half(l, i, u1, u2) : ggx_sample_half(l, i, u1, u2);
}
static ovec refl_fresnel(const orc_lobe *l, float c)
{
    return l->type == OL_CLEARCOAT ? fresnel_schlick(c, l->r0) : disney_fresnel(l, c);
}

/* ---- lobes, src/material/lobes.rs ----------------------------------------- */

int orc_lobe_kind(const orc_lobe *l)
{
    switch (l->type) {
    case OL_DISNEY_SPECULAR:
    case OL_CLEARCOAT: return LOBE_GLOSSY | LOBE_REFLECTION;             /* lobes.rs:186-188 */
    case OL_TRANSMISSION: return LOBE_GLOSSY | LOBE_TRANSMISSION;        /* lobes.rs:390-392 */
    case OL_PERFECT_MIRROR: return LOBE_SPECULAR | LOBE_REFLECTION;      /* lobes.rs:454-456 */
    default: return LOBE_DIFFUSE | LOBE_REFLECTION;                      /* lobes.rs:70-72 */
    }
}

/* DisneyDiffuseRefl::f, lobes.rs:94-115. */
static ovec diffuse_f(const orc_lobe *l, ovec i, ovec o)
{
    float f_in = schlick_weight(v_abs_cos_theta(i));
    float f_out = schlick_weight(v_abs_cos_theta(o));
    ovec diffuse = v_scale(l->color, ORC_FRAC_1_PI * (1.0f - 0.5f * f_in) * (1.0f - 0.5f * f_out));
    ovec hu = v_add(i, o);
    if (v_is_exactly_zero(hu)) return diffuse;
    ovec h = v_normalized(hu);
    float cos_d = v_dot(o, h);
    float r_r = 2.0f * l->roughness * cos_d * cos_d;
    ovec retro = v_scale(l->color, ORC_FRAC_1_PI * r_r * (f_out + f_in + f_out * f_in * (r_r - 1.0f)));
    ovec sheen = v_scale(l->sheen_color, schlick_weight(cos_d));
    return v_add(diffuse, v_add(retro, sheen));
}

/* StandardMicrofacetRefl::f, lobes.rs:135-148. */
static ovec microfacet_f(const orc_lobe *l, ovec i, ovec o)
{
    float cos_in = v_abs_cos_theta(i), cos_out = v_abs_cos_theta(o);
    ovec hu = v_add(i, o);
    if (v_is_exactly_zero(hu) || cos_in == 0.0f || cos_out == 0.0f) return v_zero();
    ovec h = v_normalized(hu);
    ovec fr = refl_fresnel(l, v_dot(o, h));
    float d = dist_d(l, h);
    float g = dist_g(l, i, o);
    return v_scale(v_mul(l->color, fr), d * g / (4.0f * cos_out * cos_in));
}

/* DisneySpecularTrans::f, lobes.rs:272-319. */
static ovec trans_f(const orc_lobe *l, ovec i, ovec o, int c2l)
{
    if (v_same_hemisphere(i, o)) return v_zero();
    float cos_in = i.z, cos_out = o.z;
    if (cos_in == 0.0f || cos_out == 0.0f) return v_zero();
    float eta = cos_in > 0.0f ? l->ior : 1.0f / l->ior;
    ovec hu = v_add(i, v_scale(o, eta));
    ovec h = hu.z > 0.0f ? v_normalized(hu) : v_neg(v_normalized(hu));
    float fd = fresnel_dielectric(v_dot(o, h), l->ior);
    ovec fr = v_scale(v_one(), fd);
    float d = ggx_d(l, h);
    float g = ggx_g(l, i, o);
    float sd = v_dot(i, h) + eta * v_dot(o, h);
    float factor = c2l ? 1.0f : eta;
    ovec inv = v_sub(v_one(), fr);
    return v_scale(v_mul(l->color, inv),
                   fabsf(d * g * factor * factor * fabsf(v_dot(o, h)) * fabsf(v_dot(i, h)) /
                         (cos_out * cos_in * sd * sd)));
}

ovec orc_lobe_f(const orc_lobe *l, ovec i, ovec o, int c2l)
{
    switch (l->type) {
    case OL_DISNEY_DIFFUSE: return diffuse_f(l, i, o);
    case OL_DISNEY_SPECULAR:
    case OL_CLEARCOAT: return microfacet_f(l, i, o);
    case OL_TRANSMISSION: return trans_f(l, i, o, c2l);
    case OL_PERFECT_DIFFUSE: return v_scale(v_new(0.5f, 0.5f, 0.5f), ORC_FRAC_1_PI); /* lobes.rs:411-413 */
    case OL_PERFECT_MIRROR: return v_div(v_one(), v_abs_cos_theta(o));               /* lobes.rs:432-434 */
    }
    return v_zero();
}

/* Default Lobe::pdf (lobes.rs:45-52), StandardMicrofacetRefl::pdf (:150-158),
 * DisneySpecularTrans::pdf with its unnormalised half vector (:321-349, Q22),
 * PerfectMirror::pdf (:436-438). */
float orc_lobe_pdf(const orc_lobe *l, ovec i, ovec o)
{
    switch (l->type) {
    case OL_DISNEY_SPECULAR:
    case OL_CLEARCOAT: {
        if (!v_same_hemisphere(i, o)) return 0.0f;
        ovec h = v_normalized(v_add(i, o));
        return dist_pdf(l, i, h) / (4.0f * v_dot(i, h));
    }
    case OL_TRANSMISSION: {
        if (v_same_hemisphere(i, o)) return 0.0f;
        float eta = i.z > 0.0f ? l->ior : 1.0f / l->ior;
        ovec hu = v_add(i, v_scale(o, eta));
        ovec h = hu.z > 0.0f ? hu : v_neg(v_normalized(hu));
        float sd = v_dot(i, h) + eta * v_dot(o, h);
        float dwh = fabsf((eta * eta * v_dot(o, h)) / (sd * sd));
        return ggx_pdf(l, i, h) * dwh;
    }
    case OL_PERFECT_MIRROR: return 1.0f;
    default:
        if (!v_same_hemisphere(i, o)) return 0.0f;
        return v_abs_cos_theta(o) * ORC_FRAC_1_PI;
    }
}

orc_lobe_sample orc_lobe_sample_f(const orc_lobe *l, ovec i, int c2l, float u1, float u2)
{
    orc_lobe_sample s = {v_zero(), v_zero(), 0.0f};
    switch (l->type) {
    case OL_DISNEY_SPECULAR:
    case OL_CLEARCOAT: { /* StandardMicrofacetRefl::sample_f, lobes.rs:160-184 */
        if (i.z == 0.0f) return s;
        ovec h = dist_sample_half(l, i, u1, u2);
        ovec o = v_reflect(i, h);
        if (!v_same_hemisphere(i, o)) return s;
        s.result = orc_lobe_f(l, i, o, c2l);
        s.outgoing = o;
        s.pdf = dist_pdf(l, i, h) / (4.0f * v_dot(i, h));
        return s;
    }
    case OL_TRANSMISSION: { /* DisneySpecularTrans::sample_f, lobes.rs:351-388 */
        if (i.z == 0.0f) return s;
        ovec h = ggx_sample_half(l, i, u1, u2);
        float eta = i.z > 0.0f ? 1.0f / l->ior : l->ior;
        ovec o = v_refract(i, h, eta);
        if (v_is_exactly_zero(o)) return s;
        s.result = orc_lobe_f(l, i, o, c2l);
        s.outgoing = o;
        s.pdf = orc_lobe_pdf(l, i, o);
        return s;
    }
    case OL_PERFECT_MIRROR: { /* lobes.rs:440-452 */
        ovec o = v_new(-i.x, -i.y, i.z);
        s.result = orc_lobe_f(l, i, o, c2l);
        s.outgoing = o;
        s.pdf = 1.0f;
        return s;
    }
    default: { /* default Lobe::sample_f, lobes.rs:54-68 */
        ovec o = orc_cosine_hemisphere(u1, u2, i.z < 0.0f);
        s.result = orc_lobe_f(l, i, o, c2l);
        s.outgoing = o;
        s.pdf = orc_lobe_pdf(l, i, o);
        return s;
    }
    }
}

/* Density over outgoing directions of orc_lobe_sample_f's draws from i (no
 * counterpart in the reference; the guided estimator's weights need it).  A
 * sub-probability: draws that fail (reflected below the surface, total
 * internal reflection) return no direction.
 *  - cosine-hemisphere and visible-normal GGX reflection draws: the reported
 *    pdf (orc_lobe_pdf);
 *  - GTR1 (gtr1_sample_half): t = cos^2 theta_h = 1 - (a^2)^(1-u2) / (1 - a^2)
 *    has dt/du2 = (1 - t) ln(a^2), so over the half-vector hemisphere the
 *    density is |cos theta_h| / (pi |ln a^2| sin^2 theta_h) for
 *    sin^2 theta_h >= a^2 / (1 - a^2) and zero nearer the normal; reflection
 *    divides by 4 i.h;
 *  - transmission: refract(i, h, eta') with i.h > 0 gives o.h < 0 and
 *    i + eta o parallel to h (eta = 1 / eta'), so h = +-normalize(i + eta o)
 *    with the sign that makes i.h > 0, valid when o.h < 0 and h lies in i's
 *    hemisphere; the density is the visible-normal density of h times
 *    eta^2 |o.h| / (i.h + eta o.h)^2;
 *  - the mirror is a delta: no density. */
float orc_lobe_density(const orc_lobe *l, ovec i, ovec o)
{
    switch (l->type) {
    case OL_CLEARCOAT: {
        if (!v_same_hemisphere(i, o)) return 0.0f;
        ovec h = v_normalized(v_add(i, o));
        float a2 = l->alpha * l->alpha;
        float sin2 = fmaxf(0.0f, 1.0f - h.z * h.z);
        if (!(sin2 >= a2 / (1.0f - a2))) return 0.0f;
        return fabsf(h.z) / ((ORC_PI * fabsf(logf(a2)) * sin2) * (4.0f * v_dot(i, h)));
    }
    case OL_TRANSMISSION: {
        if (i.z == 0.0f) return 0.0f;
        float eta = i.z > 0.0f ? l->ior : 1.0f / l->ior;
        ovec hu = v_add(i, v_scale(o, eta));
        if (v_is_exactly_zero(hu)) return 0.0f;
        ovec h = v_normalized(hu);
        if (v_dot(i, h) < 0.0f) h = v_neg(h);
        float ih = v_dot(i, h), oh = v_dot(o, h);
        if (!(ih > 0.0f && oh < 0.0f && v_same_hemisphere(h, i))) return 0.0f;
        float sd = ih + eta * oh;
        return ggx_pdf(l, i, h) * fabsf((eta * eta * oh) / (sd * sd));
    }
    case OL_PERFECT_MIRROR: return 0.0f;
    default: return orc_lobe_pdf(l, i, o);
    }
}

/* Density over world directions of orc_sample_world's draws: the mean of the
 * lobes' orc_lobe_density. */
float orc_sample_density_world(const orc_material *m, ovec wi, ovec wo, const orc_frame *f)
{
    if (m->nlobes == 0) return 0.0f;
    ovec il = v_world_to_local(wi, f->tangent, f->binormal, f->normal);
    ovec ol = v_world_to_local(wo, f->tangent, f->binormal, f->normal);
    float p = 0.0f;
    for (int k = 0; k < m->nlobes; ++k) p += orc_lobe_density(&m->lobes[k], il, ol);
    return p / (float)m->nlobes;
}

/* ---- material, src/material/material.rs ---------------------------------- */

static ovec to_local(ovec v, const orc_frame *f)
{
    return v_world_to_local(v, f->tangent, f->binormal, f->normal);
}

/* Material::f_world, material.rs:66-93. */
ovec orc_f_world(const orc_material *m, ovec wi, ovec wo, const orc_frame *f, int c2l)
{
    ovec il = to_local(wi, f), ol = to_local(wo, f);
    int reflect = (v_dot(wi, f->geom_normal) * v_dot(wo, f->geom_normal)) > 0.0f;
    ovec r = v_zero();
    for (int k = 0; k < m->nlobes; ++k) {
        int kind = orc_lobe_kind(&m->lobes[k]);
        if ((reflect && (kind & LOBE_REFLECTION)) || (!reflect && (kind & LOBE_TRANSMISSION)))
            r = v_add(r, orc_lobe_f(&m->lobes[k], il, ol, c2l));
    }
    return r;
}

/* Material::pdf_world, material.rs:95-119. */
float orc_pdf_world(const orc_material *m, ovec wi, ovec wo, const orc_frame *f)
{
    if (m->nlobes == 0) return 0.0f;
    ovec il = to_local(wi, f), ol = to_local(wo, f);
    if (il.z == 0.0f) return 0.0f;
    float pdf = 0.0f;
    for (int k = 0; k < m->nlobes; ++k) pdf += orc_lobe_pdf(&m->lobes[k], il, ol);
    return pdf / (float)m->nlobes;
}

/* Material::light_world and DiffuseAreaLight::l_world, material.rs:123-134,
 * src/material/lights.rs:12-21. */
ovec orc_light_world(const orc_material *m, ovec wi, const orc_frame *f)
{
    if (!m->has_light) return v_zero();
    return v_dot(wi, f->geom_normal) > 0.0f ? m->light_color : v_zero();
}

/* Material::sample_world, material.rs:140-235, with the lobe index drawn by
 * multiply-shift from u_lobe (Range::new(0, n), replaced). */
orc_mat_sample orc_sample_world(const orc_material *m, ovec wi, const orc_frame *f, int c2l,
                                uint32_t u_lobe, float u1, float u2)
{
    orc_mat_sample out;
    ovec il = to_local(wi, f);
    out.emission = orc_light_world(m, wi, f);
    if (m->nlobes == 0) {
        out.radiance = v_zero(); out.outgoing = v_zero(); out.pdf = 1.0f; out.kind = LOBE_NONE;
        return out;
    }
    int n = m->nlobes;
    int r = (int)(((uint64_t)u_lobe * (uint32_t)n) >> 32);
    const orc_lobe *lobe = &m->lobes[r];
    int kind = orc_lobe_kind(lobe);
    orc_lobe_sample s = orc_lobe_sample_f(lobe, il, c2l, u1, u2);
    ovec ow = v_local_to_world(s.outgoing, f->tangent, f->binormal, f->normal);
    ovec radiance = s.result;
    float pdf = s.pdf;
    if (!(kind & LOBE_SPECULAR)) {
        for (int k = 0; k < n; ++k)
            if (k != r) pdf += orc_lobe_pdf(&m->lobes[k], il, s.outgoing);
    }
    pdf /= (float)n;
    if (!(kind & LOBE_SPECULAR)) {
        int reflect = (v_dot(wi, f->geom_normal) * v_dot(ow, f->geom_normal)) > 0.0f;
        for (int k = 0; k < n; ++k) {
            /* Q16: the sampled lobe's kind gates every other lobe's f */
            if (k != r && ((reflect && (kind & LOBE_REFLECTION)) || (!reflect && (kind & LOBE_TRANSMISSION))))
                radiance = v_add(radiance, orc_lobe_f(&m->lobes[k], il, s.outgoing, c2l));
        }
    }
    if (pdf == 0.0f) {
        out.radiance = v_zero(); out.outgoing = ow; out.pdf = 1.0f; out.kind = LOBE_NONE;
        return out;
    }
    out.radiance = radiance; out.outgoing = ow; out.pdf = pdf; out.kind = kind;
    return out;
}

/* Material::count_lobes, material.rs:245-253. */
int orc_count_lobes(const orc_material *m, int mask)
{
    int c = 0;
    for (int k = 0; k < m->nlobes; ++k)
        if (orc_lobe_kind(&m->lobes[k]) & mask) c++;
    return c;
}
