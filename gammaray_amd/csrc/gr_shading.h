// gr_shading.h -- device-side BSDF evaluation: samplers, Fresnel and microfacet
// models, the Disney lobes and the material dispatch.
//
// Formulas and f32 operation order follow src/material/{lobes,util,material}.rs
// and src/core/{random,vector}.rs; the kernels are compiled without FMA
// contraction so the arithmetic matches the reference's own rounding (the
// transcendental calls -- sin, cos, log, pow -- are the device library's).
#pragma once
#pragma clang fp contract(off)

#include "gr_math.h"

namespace gr {

// LobeKind bits, src/material/lobes.rs:25-40.
enum : uint32_t {
    kLobeNone = 0,
    kLobeDiffuse = 1,
    kLobeSpecular = 2,
    kLobeGlossy = 4,
    kLobeReflection = 8,
    kLobeTransmission = 16
};

enum LobeType : uint32_t {
    kLobeDisneyDiffuse = 1,
    kLobeDisneySpecular = 2,
    kLobeClearcoat = 3,
    kLobeSpecularTrans = 4,
    kLobePerfectDiffuse = 5,
    kLobePerfectMirror = 6
};

// Parameter slots per lobe type:
//   Disney diffuse  v[0..2] colour, v[3..5] sheen colour, v[6] roughness
//   Disney specular v[0] ax, v[1] ay, v[2..4] Fresnel base colour,
//                   v[5..7] Fresnel specular colour, v[8] ior, v[9] metallic,
//                   v[10..12] lobe colour
//   clearcoat       v[0] alpha, v[1..3] r0, v[4..6] colour
//   transmission    v[0] ax, v[1] ay, v[2] ior, v[3..5] colour
struct GpuLobe {
    uint32_t type;
    float v[15];
};

struct GpuMaterial {
    uint32_t nlobes, has_light;
    float roughness;  // the builder's roughness (guide input): Disney roughness, 1 diffuse, 0 mirror
    uint32_t pad1;
    float light[4];
    float display[4];  // Material::display_color (material.rs:61-63): incandescence, one, or base colour
    GpuLobe lobes[4];
};
static_assert(sizeof(GpuMaterial) == 304, "material layout");

// Build the lobe table the reference's constructors and Disney builder make
// (src/material/material.rs:27-53,292-333).  Host side.
GpuMaterial build_material(const float rec[GR_MAT_FLOATS]);

struct Frame {
    vec3 n, t, b, gn;  // shading normal, tangent, binormal, geometric normal
};

namespace dev {

GR_HD float clamp_unit(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }
GR_HD float lerp(float x, float y, float a) { return x + a * (y - x); }
GR_HD vec3 lerp3(vec3 a, vec3 b, float t) { return mk(lerp(a.x, b.x, t), lerp(a.y, b.y, t), lerp(a.z, b.z, t)); }
GR_HD bool is_zero(vec3 v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; }
GR_HD bool nearly_zero3(vec3 v) { return nearly_zero(dot(v, v)); }
GR_HD bool same_hemi(vec3 a, vec3 b) { return a.z * b.z > 0.0f; }
GR_HD float sin2_theta(vec3 v) { return fmaxf(0.0f, 1.0f - v.z * v.z); }
GR_HD float sin_theta(vec3 v) { return sqrtf(sin2_theta(v)); }
GR_HD float cos_phi(vec3 v) {
    const float s = sin_theta(v);
    if (s == 0.0f) return 1.0f;
    const float c = v.x / s;
    return c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
}
GR_HD float sin_phi(vec3 v) {
    const float s = sin_theta(v);
    if (s == 0.0f) return 0.0f;
    const float c = v.y / s;
    return c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
}
GR_HD float luminance(vec3 c) { return 0.21f * c.x + 0.71f * c.y + 0.08f * c.z; }
GR_HD vec3 reflect(vec3 v, vec3 n) {
    const float k = 2.0f * dot(n, v);
    return mk(n.x * k - v.x, n.y * k - v.y, n.z * k - v.z);
}
GR_HD vec3 refract(vec3 v, vec3 n, float eta) {
    const float ci = dot(n, v);
    const float s2i = fmaxf(0.0f, 1.0f - ci * ci);
    const float s2t = eta * eta * s2i;
    if (s2t >= 1.0f) return mk(0.0f, 0.0f, 0.0f);
    const float ct = sqrtf(1.0f - s2t);
    return v * -eta + n * (eta * ci - ct);
}
GR_HD void coord_system(vec3 v, vec3* v2, vec3* v3) {
    if (fabsf(v.x) > fabsf(v.y)) {
        const float il = 1.0f / sqrtf(v.x * v.x + v.z * v.z);
        *v2 = mk(-v.z * il, 0.0f, v.x * il);
    } else {
        const float il = 1.0f / sqrtf(v.y * v.y + v.z * v.z);
        *v2 = mk(0.0f, v.z * il, -v.y * il);
    }
    *v3 = cross(v, *v2);
}
GR_HD vec3 to_local(vec3 v, const Frame& f) { return mk(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
GR_HD vec3 to_world(vec3 v, const Frame& f) {
    return mk(f.t.x * v.x + f.b.x * v.y + f.n.x * v.z, f.t.y * v.x + f.b.y * v.y + f.n.y * v.z,
              f.t.z * v.x + f.b.z * v.y + f.n.z * v.z);
}

// ---- samplers (src/core/random.rs) ----------------------------------------
// AreaSampleDisk with its Range(-1, 1) draws as -1 + 2u.
__device__ __forceinline__ void disk(float u1, float u2, float* x, float* y) {
    const float sx = -1.0f + 2.0f * u1, sy = -1.0f + 2.0f * u2;
    if (sx == 0.0f && sy == 0.0f) {
        *x = 0.0f;
        *y = 0.0f;
        return;
    }
    float r, th;
    if (fabsf(sx) > fabsf(sy)) {
        r = sx;
        th = 0.785398163397448309616f * (sy / sx);
    } else {
        r = sy;
        th = 1.57079632679489661923f - 0.785398163397448309616f * (sx / sy);
    }
    *x = r * cosf(th);
    *y = r * sinf(th);
}
__device__ __forceinline__ vec3 cosine_hemisphere(float u1, float u2, bool flipped) {
    float x, y;
    disk(u1, u2, &x, &y);
    const float z = sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
    return flipped ? mk(x, y, -1.0f * z) : mk(x, y, z);
}
// Uniform sphere by inversion (the reference normalises three Gaussians,
// src/core/random.rs:131-150; same distribution).
__device__ __forceinline__ vec3 uniform_sphere(float u1, float u2) {
    const float z = 1.0f - 2.0f * u1;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = (2.0f * kPi) * u2;
    return mk(r * cosf(phi), r * sinf(phi), z);
}

// ---- Fresnel (src/material/util.rs:8-110) ----------------------------------
__device__ __forceinline__ float schlick_w(float c) {
    const float x = clamp_unit(1.0f - c);
    return x * x * x * x * x;
}
__device__ __forceinline__ vec3 schlick(float c, vec3 r0) { return lerp3(r0, mk(1.0f, 1.0f, 1.0f), schlick_w(c)); }
__device__ __forceinline__ float dielectric(float cos_in, float ior) {
    float ei, et, c;
    if (cos_in > 0.0f) {
        ei = 1.0f;
        et = ior;
        c = clamp_unit(cos_in);
    } else {
        ei = ior;
        et = 1.0f;
        c = clamp_unit(-cos_in);
    }
    const float si = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    const float st = ei / et * si;
    if (st >= 1.0f) return 1.0f;
    const float ct = sqrtf(fmaxf(0.0f, 1.0f - st * st));
    const float rpa = ((et * c) - (ei * ct)) / ((et * c) + (ei * ct));
    const float rpe = ((ei * c) - (et * ct)) / ((ei * c) + (et * ct));
    return (rpa * rpa + rpe * rpe) / 2.0f;
}

// ---- GGX (util.rs:119-251) ---------------------------------------------------
__device__ __forceinline__ float ggx_lambda(float ax, float ay, vec3 v) {
    const float at = fabsf(sin_theta(v) / v.z);
    if (!isfinite(at)) return 0.0f;
    const float cp = cos_phi(v), sp = sin_phi(v);
    const float alpha = sqrtf(cp * cp * (ax * ax) + sp * sp * (ay * ay));
    const float a2t2 = (alpha * at) * (alpha * at);
    return (-1.0f + sqrtf(1.0f + a2t2)) * 0.5f;
}
__device__ __forceinline__ float ggx_d(float ax, float ay, vec3 h) {
    const float t2 = sin2_theta(h) / (h.z * h.z);
    if (!isfinite(t2)) return 0.0f;
    const float c4 = h.z * h.z * (h.z * h.z);
    const float cp = cos_phi(h), sp = sin_phi(h);
    const float e = (cp * cp / (ax * ax) + sp * sp / (ay * ay)) * t2;
    return 1.0f / (kPi * ax * ay * c4 * (1.0f + e) * (1.0f + e));
}
__device__ __forceinline__ float ggx_g(float ax, float ay, vec3 i, vec3 o) {
    return 1.0f / (1.0f + ggx_lambda(ax, ay, i) + ggx_lambda(ax, ay, o));
}
__device__ __forceinline__ float ggx_pdf(float ax, float ay, vec3 i, vec3 h) {
    if (i.z == 0.0f) return 0.0f;
    const float g1 = 1.0f / (1.0f + ggx_lambda(ax, ay, i));
    return ggx_d(ax, ay, h) * g1 * fabsf(dot(i, h)) / fabsf(i.z);
}
// Heitz's visible-normal sampling with PBRT's sample11 fit (util.rs:152-240).
__device__ __forceinline__ vec3 ggx_sample_half(float ax, float ay, vec3 i, float u1, float u2) {
    const bool flip = i.z < 0.0f;
    const vec3 f = flip ? -i : i;
    const vec3 st = normalized(mk(ax * f.x, ay * f.y, f.z));
    const float ct = st.z;
    float sx, sy;
    if (ct > 0.9999f) {
        const float r = sqrtf(u1 / (1.0f - u1));
        const float phi = (2.0f * kPi) * u2;
        sx = r * cosf(phi);
        sy = r * sinf(phi);
    } else {
        const float sn = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        const float tn = sn / ct;
        const float g1 = 2.0f / (1.0f + sqrtf(1.0f + tn * tn));
        const float a = 2.0f * u1 / g1 - 1.0f;
        const float tmp = fminf(1.0f / (a * a - 1.0f), 1e10f);
        const float b = tn;
        const float d = sqrtf(fmaxf(b * b * tmp * tmp - (a * a - b * b) * tmp, 0.0f));
        const float x1 = b * tmp - d;
        const float x2 = b * tmp + d;
        sx = (a < 0.0f || x2 > 1.0f / tn) ? x1 : x2;
        float s, u;
        if (u2 > 0.5f) {
            s = 1.0f;
            u = 2.0f * (u2 - 0.5f);
        } else {
            s = -1.0f;
            u = 2.0f * (0.5f - u2);
        }
        const float z = (u * (u * (u * 0.27385f - 0.73369f) + 0.46341f)) /
                        (u * (u * (u * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
        sy = s * z * sqrtf(1.0f + sx * sx);
    }
    const float cp = cos_phi(st), sp = sin_phi(st);
    const float rx = ax * (cp * sx - sp * sy);
    const float ry = ay * (sp * sx + cp * sy);
    const vec3 h = normalized(mk(-rx, -ry, 1.0f));
    return flip ? -h : h;
}

// ---- GTR1 (util.rs:255-309; its pdf omits cos(theta_h)) --------------------
__device__ __forceinline__ float gtr1_lambda(vec3 v) {
    const float ag = 0.25f;
    const float c = fabsf(v.z);
    const float a2 = ag * ag;
    const float c2 = c * c;
    return 1.0f / (c + sqrtf(a2 + c2 - (a2 * c2)));
}
__device__ __forceinline__ float gtr1_d(float alpha, vec3 h) {
    const float a2 = alpha * alpha;
    const float c = fabsf(h.z);
    return (a2 - 1.0f) / (kPi * logf(a2) * (1.0f + (a2 - 1.0f) * c * c));
}
__device__ __forceinline__ vec3 gtr1_sample_half(float alpha, vec3 i, float u1, float u2) {
    const float a2 = alpha * alpha;
    const float phi = 2.0f * kPi * u1;
    const float ct = sqrtf(clamp_unit(1.0f - powf(a2, 1.0f - u2) / (1.0f - a2)));
    const float snt = sqrtf(1.0f - ct * ct);
    const vec3 h = mk(snt * cosf(phi), snt * sinf(phi), ct);
    return same_hemi(h, i) ? h : -h;
}

// ---- lobes (src/material/lobes.rs) -----------------------------------------
__device__ __forceinline__ uint32_t lobe_kind(uint32_t t) {
    switch (t) {
        case kLobeDisneySpecular:
        case kLobeClearcoat: return kLobeGlossy | kLobeReflection;
        case kLobeSpecularTrans: return kLobeGlossy | kLobeTransmission;
        case kLobePerfectMirror: return kLobeSpecular | kLobeReflection;
        default: return kLobeDiffuse | kLobeReflection;
    }
}

__device__ __forceinline__ vec3 lv3(const GpuLobe& l, int k) { return mk(l.v[k], l.v[k + 1], l.v[k + 2]); }

// microfacet distribution of a reflection lobe (GGX or GTR1)
__device__ __forceinline__ float refl_d(const GpuLobe& l, uint32_t t, vec3 h) {
    return t == kLobeClearcoat ? gtr1_d(l.v[0], h) : ggx_d(l.v[0], l.v[1], h);
}
__device__ __forceinline__ float refl_g(const GpuLobe& l, uint32_t t, vec3 i, vec3 o) {
    if (t == kLobeClearcoat) return 1.0f / (1.0f + gtr1_lambda(i) + gtr1_lambda(o));
    return ggx_g(l.v[0], l.v[1], i, o);
}
__device__ __forceinline__ float refl_pdf_h(const GpuLobe& l, uint32_t t, vec3 i, vec3 h) {
    return t == kLobeClearcoat ? gtr1_d(l.v[0], h) : ggx_pdf(l.v[0], l.v[1], i, h);
}
__device__ __forceinline__ vec3 refl_fresnel(const GpuLobe& l, uint32_t t, float c) {
    if (t == kLobeClearcoat) return schlick(c, lv3(l, 1));
    // DisneyFresnel: lerp(spec_colour * F_dielectric, Schlick(base), metallic)
    const vec3 die = lv3(l, 5) * dielectric(c, l.v[8]);
    const vec3 con = schlick(c, lv3(l, 2));
    return lerp3(die, con, l.v[9]);
}
__device__ __forceinline__ vec3 refl_colour(const GpuLobe& l, uint32_t t) {
    return t == kLobeClearcoat ? lv3(l, 4) : lv3(l, 10);
}

__device__ __forceinline__ vec3 lobe_f(const GpuLobe& l, uint32_t t, vec3 i, vec3 o, bool c2l) {
    switch (t) {
        case kLobeDisneyDiffuse: {
            const float fi = schlick_w(fabsf(i.z));
            const float fo = schlick_w(fabsf(o.z));
            const vec3 col = lv3(l, 0);
            const vec3 diffuse = col * (kInvPi * (1.0f - 0.5f * fi) * (1.0f - 0.5f * fo));
            const vec3 hu = i + o;
            if (is_zero(hu)) return diffuse;
            const vec3 h = normalized(hu);
            const float cd = dot(o, h);
            const float rr = 2.0f * l.v[6] * cd * cd;
            const vec3 retro = col * (kInvPi * rr * (fo + fi + fo * fi * (rr - 1.0f)));
            const vec3 sheen = lv3(l, 3) * schlick_w(cd);
            return diffuse + (retro + sheen);
        }
        case kLobeDisneySpecular:
        case kLobeClearcoat: {
            const float ci = fabsf(i.z), co = fabsf(o.z);
            const vec3 hu = i + o;
            if (is_zero(hu) || ci == 0.0f || co == 0.0f) return mk(0.0f, 0.0f, 0.0f);
            const vec3 h = normalized(hu);
            const vec3 fr = refl_fresnel(l, t, dot(o, h));
            const float d = refl_d(l, t, h);
            const float g = refl_g(l, t, i, o);
            return mul(refl_colour(l, t), fr) * (d * g / (4.0f * co * ci));
        }
        case kLobeSpecularTrans: {
            if (same_hemi(i, o)) return mk(0.0f, 0.0f, 0.0f);
            const float ci = i.z, co = o.z;
            if (ci == 0.0f || co == 0.0f) return mk(0.0f, 0.0f, 0.0f);
            const float ior = l.v[2];
            const float eta = ci > 0.0f ? ior : 1.0f / ior;
            const vec3 hu = i + o * eta;
            const vec3 h = hu.z > 0.0f ? normalized(hu) : -normalized(hu);
            const float fd = dielectric(dot(o, h), ior);
            const vec3 fr = mk(1.0f, 1.0f, 1.0f) * fd;
            const float d = ggx_d(l.v[0], l.v[1], h);
            const float g = ggx_g(l.v[0], l.v[1], i, o);
            const float sd = dot(i, h) + eta * dot(o, h);
            const float factor = c2l ? 1.0f : eta;
            const vec3 inv = mk(1.0f, 1.0f, 1.0f) - fr;
            return mul(lv3(l, 3), inv) *
                   fabsf(d * g * factor * factor * fabsf(dot(o, h)) * fabsf(dot(i, h)) / (co * ci * sd * sd));
        }
        case kLobePerfectDiffuse: return mk(0.5f, 0.5f, 0.5f) * kInvPi;
        case kLobePerfectMirror: return mk(1.0f, 1.0f, 1.0f) / fabsf(o.z);
    }
    return mk(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ float lobe_pdf(const GpuLobe& l, uint32_t t, vec3 i, vec3 o) {
    switch (t) {
        case kLobeDisneySpecular:
        case kLobeClearcoat: {
            if (!same_hemi(i, o)) return 0.0f;
            const vec3 h = normalized(i + o);
            return refl_pdf_h(l, t, i, h) / (4.0f * dot(i, h));
        }
        case kLobeSpecularTrans: {
            if (same_hemi(i, o)) return 0.0f;
            const float ior = l.v[2];
            const float eta = i.z > 0.0f ? ior : 1.0f / ior;
            const vec3 hu = i + o * eta;
            const vec3 h = hu.z > 0.0f ? hu : -normalized(hu);  // unnormalised when z > 0
            const float sd = dot(i, h) + eta * dot(o, h);
            const float dwh = fabsf((eta * eta * dot(o, h)) / (sd * sd));
            return ggx_pdf(l.v[0], l.v[1], i, h) * dwh;
        }
        case kLobePerfectMirror: return 1.0f;
        default:
            if (!same_hemi(i, o)) return 0.0f;
            return fabsf(o.z) * kInvPi;
    }
}

// Density over outgoing directions o of lobe_sample's draws from i: a
// sub-probability, since draws that fail (reflected below the surface, total
// internal reflection) return no direction.
//  * Cosine-hemisphere (diffuse) and visible-normal GGX reflection draws: the
//    reported pdf.
//  * GTR1: the sampler sets cos^2 theta_h = 1 - (a^2)^(1-u) / (1 - a^2), so
//    with t = cos^2 theta_h, dt/du = (1 - t) ln(a^2): the half vector has
//    density |cos theta_h| / (pi |ln a^2| sin^2 theta_h) where
//    sin^2 theta_h >= a^2 / (1 - a^2) (draws nearer the normal do not occur;
//    the clamped ones lie on the horizon and reflect below), times the
//    reflection Jacobian 1 / (4 i.h).  The reported pdf is D(h) / (4 i.h).
//  * Transmission: refraction through a visible-normal GGX half vector h
//    with i.h > 0 gives o with o.h < 0 and i + eta o parallel to h (Snell),
//    so the density is the Walter et al. / PBRT pdf with h the normalised
//    i + eta o oriented to i.h > 0, zero unless o.h < 0 and h lies in i's
//    hemisphere -- also for o in i's own hemisphere, which grazing draws
//    reach.  The reported pdf leaves h unnormalised when its z > 0 and is
//    zero in i's hemisphere.
//  * The mirror: a delta, no density.
__device__ __forceinline__ float lobe_density(const GpuLobe& l, uint32_t t, vec3 i, vec3 o) {
    switch (t) {
        case kLobeClearcoat: {
            if (!same_hemi(i, o)) return 0.0f;
            const vec3 h = normalized(i + o);
            const float a2 = l.v[0] * l.v[0];
            const float s2 = sin2_theta(h);
            if (!(s2 >= a2 / (1.0f - a2))) return 0.0f;
            return fabsf(h.z) / ((kPi * fabsf(logf(a2)) * s2) * (4.0f * dot(i, h)));
        }
        case kLobeSpecularTrans: {
            if (i.z == 0.0f) return 0.0f;
            const float eta = i.z > 0.0f ? l.v[2] : 1.0f / l.v[2];
            const vec3 hu = i + o * eta;
            if (is_zero(hu)) return 0.0f;
            vec3 h = normalized(hu);
            if (dot(i, h) < 0.0f) h = -h;
            const float ih = dot(i, h), oh = dot(o, h);
            if (!(ih > 0.0f && oh < 0.0f && same_hemi(h, i))) return 0.0f;
            const float sd = ih + eta * oh;
            return ggx_pdf(l.v[0], l.v[1], i, h) * fabsf((eta * eta * oh) / (sd * sd));
        }
        case kLobePerfectMirror: return 0.0f;
        default: return lobe_pdf(l, t, i, o);
    }
}

struct LobeSample {
    vec3 result, outgoing;
    float pdf;
};

// The direction lobe l's sample_f draws from i with (u1, u2), zero when the
// draw fails; *h receives the reflection lobes' sampled half vector.
__device__ __forceinline__ vec3 lobe_draw(const GpuLobe& l, uint32_t t, vec3 i, float u1, float u2, vec3* h) {
    const vec3 zero = mk(0.0f, 0.0f, 0.0f);
    switch (t) {
        case kLobeDisneySpecular:
        case kLobeClearcoat: {
            if (i.z == 0.0f) return zero;
            *h = t == kLobeClearcoat ? gtr1_sample_half(l.v[0], i, u1, u2)
                                          : ggx_sample_half(l.v[0], l.v[1], i, u1, u2);
            const vec3 o = reflect(i, *h);
            return same_hemi(i, o) ? o : zero;
        }
        case kLobeSpecularTrans: {
            if (i.z == 0.0f) return zero;
            const vec3 hh = ggx_sample_half(l.v[0], l.v[1], i, u1, u2);
            const float eta = i.z > 0.0f ? 1.0f / l.v[2] : l.v[2];
            return refract(i, hh, eta);
        }
        case kLobePerfectMirror: return mk(-i.x, -i.y, i.z);
        default: return cosine_hemisphere(u1, u2, i.z < 0.0f);
    }
}

__device__ __forceinline__ LobeSample lobe_sample(const GpuLobe& l, uint32_t t, vec3 i, bool c2l, float u1, float u2) {
    LobeSample s{mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 0.0f), 0.0f};
    vec3 h = s.outgoing;
    const vec3 o = lobe_draw(l, t, i, u1, u2, &h);
    if (is_zero(o) && t != kLobeDisneyDiffuse && t != kLobePerfectDiffuse) return s;
    s.result = lobe_f(l, t, i, o, c2l);
    s.outgoing = o;
    switch (t) {
        case kLobeDisneySpecular:
        case kLobeClearcoat: s.pdf = refl_pdf_h(l, t, i, h) / (4.0f * dot(i, h)); break;
        case kLobePerfectMirror: s.pdf = 1.0f; break;
        default: s.pdf = lobe_pdf(l, t, i, o); break;
    }
    return s;
}

// lobe_f(l, i, o, c2l) (when want_f) and lobe_pdf(l, i, o) in one evaluation,
// the values of the two calls operation for operation, with the reflection
// lobes' half vector, microfacet density and incoming masking term computed
// once.
__device__ __forceinline__ void lobe_f_pdf(const GpuLobe& l, uint32_t t, vec3 i, vec3 o, bool c2l, bool want_f,
                                           vec3* f, float* pdf) {
    switch (t) {
        case kLobeDisneySpecular:
        case kLobeClearcoat: {
            *f = mk(0.0f, 0.0f, 0.0f);
            *pdf = 0.0f;
            const bool sh = same_hemi(i, o);
            if (!sh && !want_f) return;
            const vec3 hu = i + o;
            const vec3 h = normalized(hu);
            const float d = refl_d(l, t, h);
            const bool cc = t == kLobeClearcoat;
            const float lam_i = cc ? gtr1_lambda(i) : ggx_lambda(l.v[0], l.v[1], i);
            if (want_f) {
                const float ci = fabsf(i.z), co = fabsf(o.z);
                if (!(is_zero(hu) || ci == 0.0f || co == 0.0f)) {
                    const float lam_o = cc ? gtr1_lambda(o) : ggx_lambda(l.v[0], l.v[1], o);
                    const vec3 fr = refl_fresnel(l, t, dot(o, h));
                    const float g = 1.0f / (1.0f + lam_i + lam_o);
                    *f = mul(refl_colour(l, t), fr) * (d * g / (4.0f * co * ci));
                }
            }
            if (sh) {
                if (cc) {
                    *pdf = d / (4.0f * dot(i, h));
                } else {
                    const float g1 = i.z == 0.0f ? 0.0f : d * (1.0f / (1.0f + lam_i)) * fabsf(dot(i, h)) / fabsf(i.z);
                    *pdf = g1 / (4.0f * dot(i, h));
                }
            }
            return;
        }
        default:
            *f = want_f ? lobe_f(l, t, i, o, c2l) : mk(0.0f, 0.0f, 0.0f);
            *pdf = lobe_pdf(l, t, i, o);
            return;
    }
}

// The lobe functions with the type read from the lobe (a walk by position).
__device__ __forceinline__ uint32_t lobe_kind(const GpuLobe& l) { return lobe_kind(l.type); }
__device__ __forceinline__ vec3 lobe_f(const GpuLobe& l, vec3 i, vec3 o, bool c2l) { return lobe_f(l, l.type, i, o, c2l); }
__device__ __forceinline__ float lobe_pdf(const GpuLobe& l, vec3 i, vec3 o) { return lobe_pdf(l, l.type, i, o); }
__device__ __forceinline__ float lobe_density(const GpuLobe& l, vec3 i, vec3 o) { return lobe_density(l, l.type, i, o); }
__device__ __forceinline__ vec3 lobe_draw(const GpuLobe& l, vec3 i, float u1, float u2, vec3* h) {
    return lobe_draw(l, l.type, i, u1, u2, h);
}
__device__ __forceinline__ LobeSample lobe_sample(const GpuLobe& l, vec3 i, bool c2l, float u1, float u2) {
    return lobe_sample(l, l.type, i, c2l, u1, u2);
}

// Visits the lobes of m in position order, each with its type as a constant:
// fn(k, lobe, type).  The builder emits at most one lobe of each type, in the
// order (perfect diffuse | mirror | Disney diffuse), Disney specular,
// clearcoat, transmission, so walking the types in that order meets the
// positions in order (every sum keeps its operation order), while a wave runs
// each type's code once for all its lanes wherever the lobe sits in each
// lane's list.  A walk by position runs, at every position, the code of each
// type some lane holds there.
template <uint32_t T, typename Fn>
__device__ __forceinline__ void visit_lobe(const GpuMaterial& m, uint32_t& k, Fn& fn) {
    if (k < m.nlobes && m.lobes[k].type == T) {
        fn(k, m.lobes[k], T);
        ++k;
    }
}
template <typename Fn>
__device__ __forceinline__ void each_lobe(const GpuMaterial& m, Fn fn) {
    uint32_t k = 0;
    visit_lobe<kLobePerfectDiffuse>(m, k, fn);
    visit_lobe<kLobePerfectMirror>(m, k, fn);
    visit_lobe<kLobeDisneyDiffuse>(m, k, fn);
    visit_lobe<kLobeDisneySpecular>(m, k, fn);
    visit_lobe<kLobeClearcoat>(m, k, fn);
    visit_lobe<kLobeSpecularTrans>(m, k, fn);
}

// ---- material (src/material/material.rs) ------------------------------------
__device__ __forceinline__ vec3 light_world(const GpuMaterial& m, vec3 wi, const Frame& f) {
    if (!m.has_light) return mk(0.0f, 0.0f, 0.0f);
    return dot(wi, f.gn) > 0.0f ? mk(m.light[0], m.light[1], m.light[2]) : mk(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ vec3 f_world(const GpuMaterial& m, vec3 wi, vec3 wo, const Frame& f, bool c2l) {
    const vec3 il = to_local(wi, f), ol = to_local(wo, f);
    const bool refl = (dot(wi, f.gn) * dot(wo, f.gn)) > 0.0f;
    vec3 r = mk(0.0f, 0.0f, 0.0f);
    each_lobe(m, [&](uint32_t, const GpuLobe& l, uint32_t t) {
        const uint32_t kind = lobe_kind(t);
        if ((refl && (kind & kLobeReflection)) || (!refl && (kind & kLobeTransmission))) r = r + lobe_f(l, t, il, ol, c2l);
    });
    return r;
}

__device__ __forceinline__ float pdf_world(const GpuMaterial& m, vec3 wi, vec3 wo, const Frame& f) {
    if (m.nlobes == 0) return 0.0f;
    const vec3 il = to_local(wi, f), ol = to_local(wo, f);
    if (il.z == 0.0f) return 0.0f;
    float p = 0.0f;
    each_lobe(m, [&](uint32_t, const GpuLobe& l, uint32_t t) { p += lobe_pdf(l, t, il, ol); });
    return p / (float)m.nlobes;
}

// Density over world directions of sample_world's draws: the mean over the
// lobes of lobe_density (a mirror lobe's draws are an atom, not counted).
__device__ __forceinline__ float sample_density_world(const GpuMaterial& m, vec3 wi, vec3 wo, const Frame& f) {
    if (m.nlobes == 0) return 0.0f;
    const vec3 il = to_local(wi, f), ol = to_local(wo, f);
    float p = 0.0f;
    each_lobe(m, [&](uint32_t, const GpuLobe& l, uint32_t t) { p += lobe_density(l, t, il, ol); });
    return p / (float)m.nlobes;
}

// f_world(m, wi, wo, f, c2l), pdf_world(m, wi, wo, f) and pdf_world(m, wo, wi, f)
// in one pass -- the values of the three calls, operation for operation, with
// the terms they share evaluated once per lobe: the half vector, the
// microfacet density D(h) and the masking terms of both directions
// (BDPT's MIS weights need both densities at each joined vertex).
struct FPdf {
    vec3 f;
    float pdf_fwd, pdf_rev;
};
__device__ __forceinline__ FPdf f_pdf_world(const GpuMaterial& m, vec3 wi, vec3 wo, const Frame& f, bool c2l) {
    FPdf r;
    r.f = mk(0.0f, 0.0f, 0.0f);
    r.pdf_fwd = r.pdf_rev = 0.0f;
    const vec3 il = to_local(wi, f), ol = to_local(wo, f);
    const bool refl = (dot(wi, f.gn) * dot(wo, f.gn)) > 0.0f;
    float pf = 0.0f, pr = 0.0f;
    each_lobe(m, [&](uint32_t, const GpuLobe& l, uint32_t t) {
        const uint32_t kind = lobe_kind(t);
        const bool with_f = (refl && (kind & kLobeReflection)) || (!refl && (kind & kLobeTransmission));
        switch (t) {
            case kLobeDisneySpecular:
            case kLobeClearcoat: {
                const vec3 hu = il + ol;
                const bool sh = same_hemi(il, ol);
                if (!sh && !with_f) break;
                const vec3 h = normalized(hu);
                const float d = refl_d(l, t, h);
                float lam_i = 0.0f, lam_o = 0.0f;  // masking terms, GGX or GTR1
                if (t == kLobeClearcoat) {
                    lam_i = gtr1_lambda(il);
                    lam_o = gtr1_lambda(ol);
                } else {
                    lam_i = ggx_lambda(l.v[0], l.v[1], il);
                    lam_o = ggx_lambda(l.v[0], l.v[1], ol);
                }
                if (with_f) {
                    const float ci = fabsf(il.z), co = fabsf(ol.z);
                    if (!(is_zero(hu) || ci == 0.0f || co == 0.0f)) {
                        const vec3 fr = refl_fresnel(l, t, dot(ol, h));
                        const float g = 1.0f / (1.0f + lam_i + lam_o);
                        r.f = r.f + mul(refl_colour(l, t), fr) * (d * g / (4.0f * co * ci));
                    }
                }
                if (sh) {
                    if (t == kLobeClearcoat) {
                        pf += d / (4.0f * dot(il, h));
                        pr += d / (4.0f * dot(ol, h));
                    } else {
                        const float gi = il.z == 0.0f ? 0.0f : d * (1.0f / (1.0f + lam_i)) * fabsf(dot(il, h)) / fabsf(il.z);
                        const float go = ol.z == 0.0f ? 0.0f : d * (1.0f / (1.0f + lam_o)) * fabsf(dot(ol, h)) / fabsf(ol.z);
                        pf += gi / (4.0f * dot(il, h));
                        pr += go / (4.0f * dot(ol, h));
                    }
                }
                break;
            }
            case kLobeSpecularTrans:
                if (with_f) r.f = r.f + lobe_f(l, t, il, ol, c2l);
                pf += lobe_pdf(l, t, il, ol);
                pr += lobe_pdf(l, t, ol, il);
                break;
            case kLobePerfectMirror:
                if (with_f) r.f = r.f + lobe_f(l, t, il, ol, c2l);
                pf += 1.0f;
                pr += 1.0f;
                break;
            default:  // Disney diffuse, perfect diffuse: cosine densities
                if (with_f) r.f = r.f + lobe_f(l, t, il, ol, c2l);
                if (same_hemi(il, ol)) {
                    pf += fabsf(ol.z) * kInvPi;
                    pr += fabsf(il.z) * kInvPi;
                }
                break;
        }
    });
    if (m.nlobes != 0) {
        if (il.z != 0.0f) r.pdf_fwd = pf / (float)m.nlobes;
        if (ol.z != 0.0f) r.pdf_rev = pr / (float)m.nlobes;
    }
    return r;
}

__device__ __forceinline__ uint32_t count_lobes(const GpuMaterial& m, uint32_t mask) {
    uint32_t c = 0;
    for (uint32_t k = 0; k < m.nlobes; ++k)
        if (lobe_kind(m.lobes[k]) & mask) ++c;
    return c;
}

struct MatSample {
    vec3 emission, radiance, outgoing;
    float pdf;
    uint32_t kind;
};

// Material::sample_world after its lobe choice: lobe r (< m.nlobes) sampled
// with (u1, u2), the other lobes' densities and values added as the
// reference does.
__device__ __forceinline__ MatSample sample_world_lobe(const GpuMaterial& m, vec3 wi, const Frame& f, bool c2l,
                                                       uint32_t r, float u1, float u2) {
    MatSample out;
    const vec3 il = to_local(wi, f);
    out.emission = light_world(m, wi, f);
    if (m.nlobes == 0) {
        out.radiance = out.outgoing = mk(0.0f, 0.0f, 0.0f);
        out.pdf = 1.0f;
        out.kind = kLobeNone;
        return out;
    }
    const uint32_t n = m.nlobes;
    const GpuLobe& lobe = m.lobes[r];
    const uint32_t kind = lobe_kind(lobe);
    const LobeSample s = lobe_sample(lobe, il, c2l, u1, u2);
    const vec3 ow = to_world(s.outgoing, f);
    vec3 radiance = s.result;
    float pdf = s.pdf;
    if (!(kind & kLobeSpecular)) {
        const bool refl = (dot(wi, f.gn) * dot(ow, f.gn)) > 0.0f;
        // the sampled lobe's kind gates the other lobes' f (material.rs:203-210)
        const bool take = (refl && (kind & kLobeReflection)) || (!refl && (kind & kLobeTransmission));
        // the other lobes' densities and (when taken) values, each summed in lobe order
        each_lobe(m, [&](uint32_t k, const GpuLobe& l, uint32_t t) {
            if (k == r) return;
            vec3 fk;
            float pk;
            lobe_f_pdf(l, t, il, s.outgoing, c2l, take, &fk, &pk);
            pdf += pk;
            if (take) radiance = radiance + fk;
        });
    }
    pdf /= (float)n;
    if (pdf == 0.0f) {
        out.radiance = mk(0.0f, 0.0f, 0.0f);
        out.outgoing = ow;
        out.pdf = 1.0f;
        out.kind = kLobeNone;
        return out;
    }
    out.radiance = radiance;
    out.outgoing = ow;
    out.pdf = pdf;
    out.kind = kind;
    return out;
}

// Material::sample_world; the lobe index is the multiply-shift of u_lobe.
__device__ __forceinline__ MatSample sample_world(const GpuMaterial& m, vec3 wi, const Frame& f, bool c2l,
                                                  uint32_t u_lobe, float u1, float u2) {
    const uint32_t r = (uint32_t)(((uint64_t)u_lobe * m.nlobes) >> 32);
    return sample_world_lobe(m, wi, f, c2l, r, u1, u2);
}

}  // namespace dev
}  // namespace gr
