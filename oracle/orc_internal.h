/*
 * orc_internal.h -- shared types and f32 vector math of the CPU oracle
 * (test infrastructure only; see gr_oracle.h).
 *
 * Vector operations restate src/core/vector.rs and src/core/math.rs of the
 * reference with the same evaluation order (no FMA contraction).
 */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <float.h>
#include <math.h>
#include <stdint.h>

#define ORC_NONE 0xffffffffu
#define ORC_RAY_PUSH_DIST 1.0e-3f /* src/core/ray.rs:9 */
#define ORC_PI 3.14159265358979323846f
#define ORC_FRAC_1_PI 0.318309886183790671538f
#define ORC_FRAC_PI_2 1.57079632679489661923f
#define ORC_FRAC_PI_4 0.785398163397448309616f
#define ORC_TWO_PI (2.0f * ORC_PI) /* src/core/math.rs:3 */

typedef struct { float x, y, z; } ovec;
typedef struct { ovec min, max; } obbox;

/* src/core/vector.rs:14-285 */
static inline ovec v_new(float x, float y, float z) { ovec v = {x, y, z}; return v; }
static inline ovec v_zero(void) { return v_new(0.0f, 0.0f, 0.0f); }
static inline ovec v_one(void) { return v_new(1.0f, 1.0f, 1.0f); }
static inline ovec v_add(ovec a, ovec b) { return v_new(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline ovec v_sub(ovec a, ovec b) { return v_new(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline ovec v_scale(ovec a, float k) { return v_new(a.x * k, a.y * k, a.z * k); }
static inline ovec v_div(ovec a, float k) { return v_new(a.x / k, a.y / k, a.z / k); }
static inline ovec v_neg(ovec a) { return v_new(-a.x, -a.y, -a.z); }
static inline ovec v_mul(ovec a, ovec b) { return v_new(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float v_dot(ovec a, ovec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline ovec v_cross(ovec a, ovec b)
{
    return v_new(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float v_mag(ovec a) { return sqrtf(v_dot(a, a)); }
static inline ovec v_normalized(ovec a)
{
    float len = v_mag(a);
    return v_new(a.x / len, a.y / len, a.z / len);
}
static inline int v_is_exactly_zero(ovec a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
static inline int is_nearly_zero(float x) { return fabsf(x) < FLT_EPSILON; } /* math.rs:41-43 */
static inline int is_positive(float x) { return x > FLT_EPSILON; }          /* math.rs:48-50 */
static inline int is_close(float x, float y, float e) { return fabsf(x - y) < e; }
static inline int v_is_nearly_zero(ovec a) { return is_nearly_zero(v_dot(a, a)); }
static inline int v_is_close(ovec a, ovec b, float e)
{
    return is_close(b.x, a.x, e) && is_close(b.y, a.y, e) && is_close(b.z, a.z, e);
}
static inline float clampf(float x, float a, float b) { return x < a ? a : (x > b ? b : x); }
static inline float clamp_unit(float x) { return clampf(x, 0.0f, 1.0f); }
static inline float lerpf(float x, float y, float a) { return x + a * (y - x); }
static inline float clamped_lerp(float x, float y, float a) { return lerpf(x, y, clampf(a, 0.0f, 1.0f)); }
static inline ovec v_lerp(ovec a, ovec b, float t)
{
    return v_new(lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t));
}
static inline ovec v_sqrt(ovec a) { return v_new(sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)); }
static inline float v_luminance(ovec a) { return 0.21f * a.x + 0.71f * a.y + 0.08f * a.z; }
static inline ovec v_tint(ovec a)
{
    float l = v_luminance(a);
    return l > 0.0f ? v_div(a, l) : v_one();
}
static inline int v_is_finite(ovec a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
/* coord_system, vector.rs:91-104 */
static inline void v_coord_system(ovec v, ovec *v2, ovec *v3)
{
    if (fabsf(v.x) > fabsf(v.y)) {
        float inv_len = 1.0f / sqrtf(v.x * v.x + v.z * v.z);
        *v2 = v_new(-v.z * inv_len, 0.0f, v.x * inv_len);
    } else {
        float inv_len = 1.0f / sqrtf(v.y * v.y + v.z * v.z);
        *v2 = v_new(0.0f, v.z * inv_len, -v.y * inv_len);
    }
    *v3 = v_cross(v, *v2);
}
/* world_to_local / local_to_world, vector.rs:112-131 */
static inline ovec v_world_to_local(ovec v, ovec t, ovec b, ovec n)
{
    return v_new(v_dot(v, t), v_dot(v, b), v_dot(v, n));
}
static inline ovec v_local_to_world(ovec v, ovec t, ovec b, ovec n)
{
    return v_new(t.x * v.x + b.x * v.y + n.x * v.z,
                 t.y * v.x + b.y * v.y + n.y * v.z,
                 t.z * v.x + b.z * v.y + n.z * v.z);
}
/* local-frame trig, vector.rs:137-211 */
static inline float v_cos2_theta(ovec v) { return v.z * v.z; }
static inline float v_abs_cos_theta(ovec v) { return fabsf(v.z); }
static inline float v_sin2_theta(ovec v) { return fmaxf(0.0f, 1.0f - v_cos2_theta(v)); }
static inline float v_sin_theta(ovec v) { return sqrtf(v_sin2_theta(v)); }
static inline float v_tan_theta(ovec v) { return v_sin_theta(v) / v.z; }
static inline float v_tan2_theta(ovec v) { return v_sin2_theta(v) / v_cos2_theta(v); }
static inline float v_cos_phi(ovec v)
{
    float s = v_sin_theta(v);
    return s == 0.0f ? 1.0f : clampf(v.x / s, -1.0f, 1.0f);
}
static inline float v_sin_phi(ovec v)
{
    float s = v_sin_theta(v);
    return s == 0.0f ? 0.0f : clampf(v.y / s, -1.0f, 1.0f);
}
static inline float v_cos2_phi(ovec v) { return v_cos_phi(v) * v_cos_phi(v); }
static inline float v_sin2_phi(ovec v) { return v_sin_phi(v) * v_sin_phi(v); }
static inline int v_same_hemisphere(ovec a, ovec b) { return a.z * b.z > 0.0f; }
/* reflect / refract, vector.rs:240-264 */
static inline ovec v_reflect(ovec v, ovec n)
{
    float k = 2.0f * v_dot(n, v);
    return v_new(n.x * k - v.x, n.y * k - v.y, n.z * k - v.z);
}
static inline ovec v_refract(ovec v, ovec n, float eta)
{
    float cos_in = v_dot(n, v);
    float sin2_in = fmaxf(0.0f, 1.0f - cos_in * cos_in);
    float sin2_tr = eta * eta * sin2_in;
    if (sin2_tr >= 1.0f) return v_zero();
    float cos_tr = sqrtf(1.0f - sin2_tr);
    return v_add(v_scale(v, -eta), v_scale(n, eta * cos_in - cos_tr));
}

/* ---- lobes and materials (src/material/) ---- */
enum { LOBE_NONE = 0, LOBE_DIFFUSE = 1, LOBE_SPECULAR = 2, LOBE_GLOSSY = 4,
       LOBE_REFLECTION = 8, LOBE_TRANSMISSION = 16 }; /* lobes.rs:25-40 */

enum { OL_DISNEY_DIFFUSE = 1, OL_DISNEY_SPECULAR, OL_CLEARCOAT, OL_TRANSMISSION,
       OL_PERFECT_DIFFUSE, OL_PERFECT_MIRROR };

typedef struct {
    int type;
    ovec color;       /* diffuse colour / microfacet lobe colour / trans colour */
    ovec sheen_color; /* DisneyDiffuseRefl */
    float roughness;  /* DisneyDiffuseRefl */
    float ax, ay;     /* GGX */
    float alpha;      /* GTR1 */
    ovec fr_color, fr_spec_color; /* DisneyFresnel */
    float fr_ior, fr_metallic;
    ovec r0;          /* SchlickFresnel */
    float ior;        /* transmission */
} orc_lobe;

typedef struct {
    int has_light;
    ovec light_color;
    ovec display;
    int nlobes;
    orc_lobe lobes[4];
} orc_material;

void orc_material_init(orc_material *m, const float *rec);

/* ---- prims / scene ---- */
enum { ORC_PRIM_MESH = 0, ORC_PRIM_SPHERE = 1 };

typedef struct {
    int type;
    orc_material mat;
    /* mesh */
    ovec *verts, *normals, *uvs;
    uint32_t *tris; /* 9 per tri */
    uint32_t nverts, nnormals, nuvs, ntris;
    float *area_cdf;
    float area;
    /* sphere */
    ovec center;
    float radius;
} orc_prim;

typedef struct {
    obbox bbox;
    uint32_t offset, n, axis;
} orc_node;

struct orc_scene {
    orc_prim *prims;
    uint32_t num_prims, cap_prims;
    orc_node *nodes;
    uint32_t num_nodes;
    uint32_t *comps; /* pairs (prim, comp) */
    uint32_t num_comps;
    uint32_t *lights;
    uint32_t num_lights;
    float slab_pad;
};

typedef struct { uint32_t prim, comp; float dist; } orc_hit;
typedef struct { ovec normal, tangent, binormal, geom_normal; } orc_frame;

float orc_tri_intersect(const orc_prim *m, uint32_t comp, ovec o, ovec d, float *bu, float *bv);
float orc_sphere_intersect(const orc_prim *sp, ovec o, ovec d);
int orc_intersect_one(const struct orc_scene *s, ovec o, ovec d, orc_hit *hit, uint32_t cnt[3]);
int orc_visibility_one(const struct orc_scene *s, ovec start, ovec target, uint32_t cnt[3]);
void orc_hit_frame(const struct orc_scene *s, const orc_hit *h, ovec o, ovec d, orc_frame *f);
void orc_mesh_frame(const orc_prim *m, uint32_t comp, float u, float v, float w, orc_frame *f);
void orc_sphere_frame(const orc_prim *sp, ovec pt, orc_frame *f);

#endif
