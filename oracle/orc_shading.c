/*
 * orc_shading.c -- CPU oracle: surface frames and the Disney material builder
 * (test infrastructure only; see gr_oracle.h).
 */
#include "gr_oracle.h"
#include "orc_internal.h"

#include <string.h>

/* Mesh::compute_surface_props, src/geom/mesh.rs:190-240 (including the uv
 * determinant written with '+', Q10, and dpdu built from the uv deltas). */
void orc_mesh_frame(const orc_prim *m, uint32_t comp, float u, float v, float w, orc_frame *f)
{
    const uint32_t *t = &m->tris[9 * (size_t)comp];
    ovec a = m->verts[t[0]], b = m->verts[t[1]], c = m->verts[t[2]];
    ovec edge1 = v_sub(a, c), edge2 = v_sub(b, c);
    ovec at = m->uvs[t[6]], bt = m->uvs[t[7]], ct = m->uvs[t[8]];
    ovec uv1 = v_sub(at, ct), uv2 = v_sub(bt, ct);
    ovec geom_normal = v_normalized(v_cross(edge1, edge2));
    ovec normal;
    if (t[3] == ORC_NONE) {
        normal = geom_normal;
    } else {
        ovec an = m->normals[t[3]], bn = m->normals[t[4]], cn = m->normals[t[5]];
        normal = v_normalized(v_add(v_add(v_scale(an, u), v_scale(bn, v)), v_scale(cn, w)));
    }
    float uv_det = uv1.x * uv2.y + uv1.y * uv2.x;
    ovec tangent, binormal;
    if (uv_det == 0.0f) {
        v_coord_system(normal, &tangent, &binormal);
    } else {
        float inv = 1.0f / uv_det;
        ovec dpdu = v_scale(v_sub(v_scale(uv1, uv2.y), v_scale(uv2, uv1.y)), inv);
        binormal = v_normalized(v_cross(normal, dpdu));
        tangent = v_cross(binormal, normal);
    }
    f->normal = normal; f->tangent = tangent; f->binormal = binormal; f->geom_normal = geom_normal;
}

/* Sphere::compute_surface_props, src/geom/sphere.rs:28-45. */
void orc_sphere_frame(const orc_prim *sp, ovec pt, orc_frame *f)
{
    ovec normal = v_normalized(v_sub(pt, sp->center));
    ovec tangent;
    if (is_nearly_zero(normal.x) && is_nearly_zero(normal.z)) tangent = v_new(1.0f, 0.0f, 0.0f);
    else tangent = v_normalized(v_new(-normal.z, 0.0f, normal.x));
    f->normal = normal;
    f->tangent = tangent;
    f->binormal = v_cross(normal, tangent);
    f->geom_normal = normal;
}

/* Frame of a traversal hit: the SurfaceProperties that intersect_world returns
 * with the winning distance (mesh.rs:305-307, sphere.rs:416-423). */
void orc_hit_frame(const orc_scene *s, const orc_hit *h, ovec o, ovec d, orc_frame *f)
{
    const orc_prim *p = &s->prims[h->prim];
    if (p->type == ORC_PRIM_MESH) {
        float u = 0.0f, v = 0.0f;
        orc_tri_intersect(p, h->comp, o, d, &u, &v);
        float w = 1.0f - u - v;
        orc_mesh_frame(p, h->comp, u, v, w, f);
    } else {
        ovec pt = v_add(o, v_scale(d, h->dist)); /* Ray::at, ray.rs:22-24 */
        orc_sphere_frame(p, pt, f);
    }
}

/* GgxDistribution::new, src/material/util.rs:128-133. */
static void ggx_init(orc_lobe *l, float roughness, float anisotropic)
{
    float aspect = sqrtf(1.0f - anisotropic * 0.9f);
    l->ax = fmaxf(0.001f, roughness * roughness / aspect);
    l->ay = fmaxf(0.001f, roughness * roughness * aspect);
}

/* Material constructors, src/material/material.rs:27-53, and
 * DisneyMaterialBuilder::build, material.rs:292-333 (lobe constructors at
 * src/material/lobes.rs:83-91,212-225,232-242,258-268). */
void orc_material_init(orc_material *m, const float *rec)
{
    memset(m, 0, sizeof(*m));
    int kind = (int)rec[0];
    if (kind == 1) {
        m->has_light = 1;
        m->light_color = v_new(rec[1], rec[2], rec[3]);
        m->display = m->light_color;
        return;
    }
    if (kind == 2) {
        m->display = v_one();
        m->nlobes = 1;
        m->lobes[0].type = OL_PERFECT_DIFFUSE;
        return;
    }
    if (kind == 3) {
        m->display = v_one();
        m->nlobes = 1;
        m->lobes[0].type = OL_PERFECT_MIRROR;
        return;
    }
    ovec base = v_new(rec[1], rec[2], rec[3]);
    float roughness = rec[4], anisotropic = rec[5], ior = rec[6], metallic = rec[7];
    float specular_trans = rec[8], specular_tint = rec[9], sheen = rec[10], sheen_tint = rec[11];
    float clearcoat = rec[12], clearcoat_gloss = rec[13];
    float diffuse_weight = (1.0f - metallic) * (1.0f - specular_trans);
    float trans_weight = (1.0f - metallic) * specular_trans;
    m->display = base;
    if (diffuse_weight > 0.0f) {
        orc_lobe *l = &m->lobes[m->nlobes++];
        l->type = OL_DISNEY_DIFFUSE;
        l->color = v_scale(base, diffuse_weight);
        l->sheen_color = v_scale(v_lerp(v_one(), v_tint(base), sheen_tint), sheen * diffuse_weight);
        l->roughness = roughness;
    }
    if (ior > 1.0f) {
        orc_lobe *l = &m->lobes[m->nlobes++];
        l->type = OL_DISNEY_SPECULAR;
        float ior_adj = fmaxf(ior, 1.01f);
        ggx_init(l, roughness, anisotropic);
        l->fr_color = base;
        l->fr_spec_color = v_lerp(v_one(), v_tint(base), specular_tint);
        l->fr_ior = ior_adj;
        l->fr_metallic = metallic;
        l->color = v_one();
    }
    if (clearcoat > 0.0f) {
        orc_lobe *l = &m->lobes[m->nlobes++];
        l->type = OL_CLEARCOAT;
        l->alpha = lerpf(0.1f, 0.001f, clearcoat_gloss);
        l->r0 = v_scale(v_one(), 0.04f);
        l->color = v_scale(v_one(), 0.25f * clearcoat);
    }
    if (trans_weight > 0.0f) {
        orc_lobe *l = &m->lobes[m->nlobes++];
        l->type = OL_TRANSMISSION;
        l->color = v_scale(v_sqrt(base), trans_weight);
        ggx_init(l, roughness, anisotropic);
        l->ior = fmaxf(ior, 1.01f);
    }
}
