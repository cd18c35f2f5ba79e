/*
 * gr_oracle.c -- CPU oracle (test infrastructure; see gr_oracle.h header).
 *
 * A plain-C restatement of sdao/gammaray.  Each function names the reference
 * file:line it follows.  Compile with -ffp-contract=off -fno-fast-math.
 */
#include "gr_oracle.h"
#include "orc_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Scene construction                                                        */
/* ------------------------------------------------------------------------ */

int orc_scene_create(orc_scene **out)
{
    orc_scene *s = (orc_scene *)calloc(1, sizeof(orc_scene));
    if (!s) return -1;
    *out = s;
    return 0;
}

static void free_tree(orc_scene *s)
{
    free(s->nodes); s->nodes = NULL; s->num_nodes = 0;
    free(s->comps); s->comps = NULL; s->num_comps = 0;
    free(s->lights); s->lights = NULL; s->num_lights = 0;
}

void orc_scene_destroy(orc_scene *s)
{
    if (!s) return;
    for (uint32_t i = 0; i < s->num_prims; ++i) {
        orc_prim *p = &s->prims[i];
        free(p->verts); free(p->normals); free(p->uvs); free(p->tris); free(p->area_cdf);
    }
    free(s->prims);
    free_tree(s);
    free(s);
}

static orc_prim *push_prim(orc_scene *s)
{
    if (s->num_prims == s->cap_prims) {
        uint32_t nc = s->cap_prims ? s->cap_prims * 2 : 16;
        orc_prim *np = (orc_prim *)realloc(s->prims, nc * sizeof(orc_prim));
        if (!np) return NULL;
        s->prims = np;
        s->cap_prims = nc;
    }
    orc_prim *p = &s->prims[s->num_prims];
    memset(p, 0, sizeof(*p));
    return p;
}

/* Triangle area, src/geom/mesh.rs:37-47. */
static float tri_area(const orc_prim *m, const uint32_t *t)
{
    ovec a = m->verts[t[0]], b = m->verts[t[1]], c = m->verts[t[2]];
    ovec e1 = v_sub(a, c), e2 = v_sub(b, c);
    return 0.5f * v_mag(v_cross(e1, e2));
}

int orc_scene_add_mesh(orc_scene *s, const float *verts, uint32_t nverts,
                       const float *normals, uint32_t nnormals,
                       const float *uvs, uint32_t nuvs,
                       const uint32_t *tris, uint32_t ntris, const float *mat)
{
    if (ntris == 0 || nverts == 0) return -1;
    for (uint64_t i = 0; i < (uint64_t)ntris * 9; ++i) {
        uint32_t k = (uint32_t)(i % 9), v = tris[i];
        if (k < 3 && v >= nverts) return -2;
        if (k >= 3 && k < 6 && v != ORC_NONE && v >= nnormals) return -2;
        if (k >= 6 && v != ORC_NONE && v >= nuvs) return -2;
    }
    orc_prim *p = push_prim(s);
    if (!p) return -1;
    p->type = ORC_PRIM_MESH;
    p->nverts = nverts; p->nnormals = nnormals; p->nuvs = nuvs; p->ntris = ntris;
    p->verts = (ovec *)malloc(sizeof(ovec) * nverts);
    p->normals = (ovec *)malloc(sizeof(ovec) * (nnormals ? nnormals : 1));
    p->uvs = (ovec *)malloc(sizeof(ovec) * (nuvs + 1));
    p->tris = (uint32_t *)malloc(sizeof(uint32_t) * 9 * (size_t)ntris);
    p->area_cdf = (float *)malloc(sizeof(float) * ntris);
    for (uint32_t i = 0; i < nverts; ++i) p->verts[i] = v_new(verts[3*i], verts[3*i+1], verts[3*i+2]);
    /* Zero-length normals are kept as given and triangles that reference one
     * fall back to the geometric normal (src/geom/mesh.rs:109-117,138-153). */
    for (uint32_t i = 0; i < nnormals; ++i) p->normals[i] = v_new(normals[3*i], normals[3*i+1], normals[3*i+2]);
    for (uint32_t i = 0; i < nuvs; ++i) p->uvs[i] = v_new(uvs[2*i], uvs[2*i+1], 0.0f);
    p->uvs[nuvs] = v_zero(); /* the per-triangle zero uv of src/geom/mesh.rs:133-136 */
    for (uint32_t i = 0; i < ntris; ++i) {
        uint32_t *t = &p->tris[9 * (size_t)i];
        memcpy(t, &tris[9 * (size_t)i], 9 * sizeof(uint32_t));
        int has_n = t[3] != ORC_NONE && t[4] != ORC_NONE && t[5] != ORC_NONE;
        if (has_n && (v_is_nearly_zero(p->normals[t[3]]) || v_is_nearly_zero(p->normals[t[4]]) ||
                      v_is_nearly_zero(p->normals[t[5]])))
            has_n = 0;
        if (!has_n) t[3] = t[4] = t[5] = ORC_NONE;
        if (t[6] == ORC_NONE || t[7] == ORC_NONE || t[8] == ORC_NONE) t[6] = t[7] = t[8] = nuvs;
    }
    /* Area CDF, src/geom/mesh.rs:161-170. */
    float total = 0.0f;
    for (uint32_t i = 0; i < ntris; ++i) {
        total += tri_area(p, &p->tris[9 * (size_t)i]);
        p->area_cdf[i] = total;
    }
    for (uint32_t i = 0; i < ntris; ++i) p->area_cdf[i] = p->area_cdf[i] / total;
    p->area = total;
    orc_material_init(&p->mat, mat);
    return (int)(s->num_prims++);
}

int orc_scene_add_sphere(orc_scene *s, const float *center, float radius, const float *mat)
{
    orc_prim *p = push_prim(s);
    if (!p) return -1;
    p->type = ORC_PRIM_SPHERE;
    p->center = v_new(center[0], center[1], center[2]);
    p->radius = radius;
    orc_material_init(&p->mat, mat);
    return (int)(s->num_prims++);
}

/* ------------------------------------------------------------------------ */
/* BBox helpers, src/core/bbox.rs                                            */
/* ------------------------------------------------------------------------ */

/* BBox::empty, src/core/bbox.rs:15-20 (f32::MIN == -FLT_MAX). */
static obbox bb_empty(void)
{
    obbox b = {{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    return b;
}
/* union_with, src/core/bbox.rs:26-33. */
static obbox bb_union(obbox b, ovec k)
{
    obbox r = {{fminf(b.min.x, k.x), fminf(b.min.y, k.y), fminf(b.min.z, k.z)},
               {fmaxf(b.max.x, k.x), fmaxf(b.max.y, k.y), fmaxf(b.max.z, k.z)}};
    return r;
}
/* combine_with, src/core/bbox.rs:35-46. */
static obbox bb_combine(obbox a, obbox b)
{
    obbox r = {{fminf(a.min.x, b.min.x), fminf(a.min.y, b.min.y), fminf(a.min.z, b.min.z)},
               {fmaxf(a.max.x, b.max.x), fmaxf(a.max.y, b.max.y), fmaxf(a.max.z, b.max.z)}};
    return r;
}
/* maximum_extent, src/core/bbox.rs:52-63. */
static int bb_max_extent(obbox b)
{
    ovec d = v_sub(b.max, b.min);
    if (d.x >= d.y && d.y >= d.z) return 0;
    else if (d.y >= d.z) return 1;
    return 2;
}
/* relative_offset, src/core/bbox.rs:67-71. */
static ovec bb_rel(obbox b, ovec v)
{
    ovec a = v_sub(v, b.min);
    ovec d = v_sub(b.max, b.min);
    return v_new(a.x / d.x, a.y / d.y, a.z / d.z);
}
/* surface_area (a volume: Q1), src/core/bbox.rs:73-76. */
static float bb_sa(obbox b)
{
    ovec d = v_sub(b.max, b.min);
    return d.x * d.y * d.z;
}

/* gamma(3) of src/core/math.rs:152-154, folded as the reference evaluates it:
 * (3*eps) / (1 - 3*eps), then 1 + 2*gamma. */
static float slab_pad(void)
{
    volatile float n = 3.0f;
    float g = (n * FLT_EPSILON) / (1.0f - (n * FLT_EPSILON));
    return 1.0f + 2.0f * g;
}

/* BBox::intersect, src/core/bbox.rs:78-113.  Returns 1 on hit and writes the
 * slab entry t (tx/ty only; the z slab never raises t_min -- Q7). */
static int bb_intersect(const obbox *b, ovec o, ovec inv, const int neg[3], float max_dist,
                        float pad)
{
    const ovec *lo_hi[2] = {&b->min, &b->max};
    float t_min = (lo_hi[neg[0]]->x - o.x) * inv.x;
    float t_max = (lo_hi[!neg[0]]->x - o.x) * inv.x;
    float ty_min = (lo_hi[neg[1]]->y - o.y) * inv.y;
    float ty_max = (lo_hi[!neg[1]]->y - o.y) * inv.y;
    t_max *= pad;
    ty_max *= pad;
    if (t_min > ty_max || ty_min > t_max) return 0;
    if (ty_min > t_min) t_min = ty_min;
    if (ty_max < t_max) t_max = ty_max;
    float tz_min = (lo_hi[neg[2]]->z - o.z) * inv.z;
    float tz_max = (lo_hi[!neg[2]]->z - o.z) * inv.z;
    tz_max *= pad;
    if (t_min > tz_max || tz_min > t_max) return 0;
    if (tz_max < t_max) t_max = tz_max;
    return t_min < max_dist && t_max > 0.0f;
}

/* BBox::intersect on n (box, ray) pairs: boxes as lo.xyz hi.xyz, rays as
 * o.xyz d.xyz, 1/d and the direction signs as Ray::compute_intersection_data
 * (src/core/ray.rs:31-42) forms them.  out[i] = 1 on a hit before max_dist[i]. */
void orc_bbox_intersect(const float *boxes, const float *rays, const float *max_dist, int64_t n,
                        uint8_t *out)
{
    const float pad = slab_pad();
    for (int64_t i = 0; i < n; ++i) {
        const float *b = boxes + 6 * i, *r = rays + 6 * i;
        const obbox bb = {{b[0], b[1], b[2]}, {b[3], b[4], b[5]}};
        const ovec o = v_new(r[0], r[1], r[2]);
        const ovec inv = v_new(1.0f / r[3], 1.0f / r[4], 1.0f / r[5]);
        const int neg[3] = {r[3] < 0.0f, r[4] < 0.0f, r[5] < 0.0f};
        out[i] = (uint8_t)bb_intersect(&bb, o, inv, neg, max_dist[i], pad);
    }
}

/* ------------------------------------------------------------------------ */
/* Prim bounds and intersection                                             */
/* ------------------------------------------------------------------------ */

/* Mesh::bbox_world, src/geom/mesh.rs:262-268; Sphere::bbox_world,
 * src/geom/sphere.rs:57-68. */
static obbox prim_bbox(const orc_prim *p, uint32_t comp)
{
    if (p->type == ORC_PRIM_MESH) {
        const uint32_t *t = &p->tris[9 * (size_t)comp];
        obbox b = bb_empty();
        b = bb_union(b, p->verts[t[0]]);
        b = bb_union(b, p->verts[t[1]]);
        b = bb_union(b, p->verts[t[2]]);
        return b;
    }
    obbox b = {{p->center.x - p->radius, p->center.y - p->radius, p->center.z - p->radius},
               {p->center.x + p->radius, p->center.y + p->radius, p->center.z + p->radius}};
    return b;
}

/* Mesh::intersect_world distance part, src/geom/mesh.rs:270-304.  Returns 0 on
 * a miss (the reference's miss sentinel), else dist; *uvw gets (u, v) for the
 * frame.  The frame itself (mesh.rs:305-306) does not influence the hit choice,
 * so it is computed once for the final hit (orc_tri_frame). */
float orc_tri_intersect(const orc_prim *m, uint32_t comp, ovec o, ovec d, float *bu, float *bv)
{
    const uint32_t *t = &m->tris[9 * (size_t)comp];
    ovec a = m->verts[t[0]], b = m->verts[t[1]], c = m->verts[t[2]];
    ovec edge1 = v_sub(a, c);
    ovec edge2 = v_sub(b, c);
    ovec p = v_cross(d, edge2);
    float det = v_dot(edge1, p);
    if (is_nearly_zero(det)) return 0.0f;
    float inv_det = 1.0f / det;
    ovec tv = v_sub(o, c);
    float u = v_dot(tv, p) * inv_det;
    if (u < 0.0f || u > 1.0f) return 0.0f;
    ovec q = v_cross(tv, edge1);
    float v = v_dot(d, q) * inv_det;
    if (v < 0.0f || (u + v) > 1.0f) return 0.0f;
    float dist = v_dot(edge2, q) * inv_det;
    if (!is_positive(dist)) return 0.0f;
    if (bu) *bu = u;
    if (bv) *bv = v;
    return dist;
}

/* Sphere::intersect_world distance part, src/geom/sphere.rs:70-101. */
float orc_sphere_intersect(const orc_prim *sp, ovec o, ovec d)
{
    ovec origin = v_sub(o, sp->center);
    ovec l = d;
    float a = v_dot(l, l);
    float b = v_dot(l, origin);
    float c = v_dot(origin, origin) - (sp->radius * sp->radius);
    float disc = (b * b) - (a * c);
    if (disc > 0.0f) {
        float sq = sqrtf(disc);
        float res_pos = -b + sq;
        float res_neg = -b - sq;
        if (is_positive(res_neg)) return res_neg;
        else if (is_positive(res_pos)) return res_pos;
    }
    return 0.0f;
}

/* ------------------------------------------------------------------------ */
/* BVH build, src/geom/bvh.rs:182-382 and src/geom/util.rs:4-46             */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t prim, comp;
    obbox bbox;
    ovec centroid;
} comp_info; /* BvhComponentInfo, src/geom/bvh.rs:15-34 */

typedef struct {
    obbox bbox;
    uint32_t child[2];
    uint32_t axis;
    uint32_t first;
    uint32_t n; /* 0 = interior */
} build_node; /* BvhBuildNode, src/geom/bvh.rs:48-85 */

typedef struct {
    build_node *nodes;
    size_t n, cap;
    uint32_t *ordered; /* pairs */
    size_t n_ordered;
} build_ctx;

static uint32_t arena_push(build_ctx *bc, build_node nd)
{
    if (bc->n == bc->cap) {
        bc->cap = bc->cap ? bc->cap * 2 : 1024;
        bc->nodes = (build_node *)realloc(bc->nodes, bc->cap * sizeof(build_node));
    }
    bc->nodes[bc->n] = nd;
    return (uint32_t)(bc->n++);
}

static float vec_get(ovec v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

/* util::partition, src/geom/util.rs:4-17 (the predicate is passed as a
 * comparison against a pivot centroid or a bucket rule). */
typedef int (*pred_fn)(const comp_info *ci, const void *ctx);

static size_t util_partition(comp_info *sl, size_t len, pred_fn pred, const void *ctx)
{
    size_t cursor = 0;
    for (size_t i = 0; i < len; ++i) {
        if (pred(&sl[i], ctx)) {
            comp_info tmp = sl[i]; sl[i] = sl[cursor]; sl[cursor] = tmp;
            cursor += 1;
        }
    }
    return cursor;
}

typedef struct { int dim; float pivot; } less_ctx;
static int pred_less(const comp_info *ci, const void *vctx)
{
    const less_ctx *c = (const less_ctx *)vctx;
    return vec_get(ci->centroid, c->dim) < c->pivot;
}

/* util::nth_element, src/geom/util.rs:23-46, including the left-recursion range
 * 0..(pivot-1) that drops one element (Q4).  Tail recursion written as a loop. */
static void util_nth_element(comp_info *sl, size_t len, size_t nth, int dim)
{
    for (;;) {
        if (len < 2) return;
        size_t pivot = len / 2;
        comp_info tmp = sl[pivot]; sl[pivot] = sl[len - 1]; sl[len - 1] = tmp;
        less_ctx lc = {dim, vec_get(sl[len - 1].centroid, dim)};
        size_t fpi = util_partition(sl, len - 1, pred_less, &lc);
        tmp = sl[len - 1]; sl[len - 1] = sl[fpi]; sl[fpi] = tmp;
        if (nth < fpi) {
            len = fpi - 1;
        } else if (nth > fpi) {
            sl = sl + fpi + 1;
            len = len - (fpi + 1);
            nth = nth - fpi - 1;
        } else {
            return;
        }
    }
}

#define NUM_BUCKETS 12

/* (NUM_BUCKETS as f32 * rel[dim]) as usize, clamped (Q6): Rust's float->usize
 * cast saturates (NaN -> 0). */
static size_t bucket_of(obbox cb, ovec centroid, int dim)
{
    float r = vec_get(bb_rel(cb, centroid), dim);
    float f = (float)NUM_BUCKETS * r;
    size_t b;
    if (!(f > 0.0f)) b = 0;                 /* NaN and negatives */
    else if (f >= 18446744073709551616.0f) b = SIZE_MAX;
    else b = (size_t)f;
    if (b > NUM_BUCKETS - 1) b = NUM_BUCKETS - 1;
    return b;
}

typedef struct { obbox cb; int dim; size_t split; } bucket_ctx;
static int pred_bucket(const comp_info *ci, const void *vctx)
{
    const bucket_ctx *c = (const bucket_ctx *)vctx;
    return bucket_of(c->cb, ci->centroid, c->dim) <= c->split;
}

/* Bvh::recurse_build, src/geom/bvh.rs:182-308. */
static uint32_t recurse_build(build_ctx *bc, comp_info *ci, size_t n)
{
    obbox bbox = bb_empty();
    for (size_t i = 0; i < n; ++i) bbox = bb_combine(bbox, ci[i].bbox);

    if (n == 1) {
        build_node nd = {bbox, {0, 0}, 0, (uint32_t)bc->n_ordered, 1};
        bc->ordered[2 * bc->n_ordered] = ci[0].prim;
        bc->ordered[2 * bc->n_ordered + 1] = ci[0].comp;
        bc->n_ordered++;
        return arena_push(bc, nd);
    }
    obbox cb = bb_empty();
    for (size_t i = 0; i < n; ++i) cb = bb_union(cb, ci[i].centroid);
    int dim = bb_max_extent(cb);
    if (vec_get(cb.min, dim) == vec_get(cb.max, dim)) {
        build_node nd = {bbox, {0, 0}, 0, (uint32_t)bc->n_ordered, (uint32_t)n};
        for (size_t i = 0; i < n; ++i) {
            bc->ordered[2 * bc->n_ordered] = ci[i].prim;
            bc->ordered[2 * bc->n_ordered + 1] = ci[i].comp;
            bc->n_ordered++;
        }
        return arena_push(bc, nd);
    }
    size_t mid;
    if (n <= 4) {
        mid = n / 2;
        util_nth_element(ci, n, mid, dim);
    } else {
        size_t count[NUM_BUCKETS];
        obbox bb[NUM_BUCKETS];
        for (int i = 0; i < NUM_BUCKETS; ++i) { count[i] = 0; bb[i] = bb_empty(); }
        for (size_t i = 0; i < n; ++i) {
            size_t b = bucket_of(cb, ci[i].centroid, dim);
            count[b] += 1;
            bb[b] = bb_combine(bb[b], ci[i].bbox);
        }
        float cost[NUM_BUCKETS - 1];
        for (int i = 0; i < NUM_BUCKETS - 1; ++i) {
            obbox b0 = bb_empty(), b1 = bb_empty();
            size_t c0 = 0, c1 = 0;
            for (int j = 0; j < i + 1; ++j) { b0 = bb_combine(b0, bb[j]); c0 += count[j]; }
            for (int j = i + 1; j < NUM_BUCKETS; ++j) { b1 = bb_combine(b1, bb[j]); c1 += count[j]; }
            cost[i] = 1.0f + ((float)c0 * bb_sa(b0) + (float)c1 * bb_sa(b1)) / bb_sa(bbox);
        }
        float min_cost = cost[0];
        size_t split = 0;
        for (int i = 1; i < NUM_BUCKETS - 1; ++i) {
            if (cost[i] < min_cost) { min_cost = cost[i]; split = (size_t)i; }
        }
        float leaf_cost = (float)n;
        if (n > 255 || min_cost < leaf_cost) {
            bucket_ctx bctx = {cb, dim, split};
            mid = util_partition(ci, n, pred_bucket, &bctx);
        } else {
            mid = n / 2;
            util_nth_element(ci, n, mid, dim);
        }
    }
    uint32_t c0 = recurse_build(bc, ci, mid);
    uint32_t c1 = recurse_build(bc, ci + mid, n - mid);
    build_node nd;
    nd.bbox = bb_combine(bc->nodes[c0].bbox, bc->nodes[c1].bbox);
    nd.child[0] = c0; nd.child[1] = c1;
    nd.axis = (uint32_t)dim; nd.first = 0; nd.n = 0;
    return arena_push(bc, nd);
}

/* Bvh::flatten_tree, src/geom/bvh.rs:310-337 (DFS: first child = i + 1). */
static uint32_t flatten(const build_ctx *bc, orc_node *out, uint32_t *count, uint32_t root)
{
    const build_node *bn = &bc->nodes[root];
    uint32_t idx = (*count)++;
    if (bn->n > 0) {
        out[idx].bbox = bn->bbox;
        out[idx].offset = bn->first;
        out[idx].n = bn->n;
        out[idx].axis = 0;
        return idx;
    }
    flatten(bc, out, count, bn->child[0]);
    uint32_t second = flatten(bc, out, count, bn->child[1]);
    out[idx].bbox = bn->bbox;
    out[idx].offset = second;
    out[idx].n = 0;
    out[idx].axis = bn->axis;
    return idx;
}

/* Bvh::build, src/geom/bvh.rs:339-382. */
int orc_scene_build(orc_scene *s)
{
    free_tree(s);
    size_t total = 0;
    for (uint32_t i = 0; i < s->num_prims; ++i)
        total += s->prims[i].type == ORC_PRIM_MESH ? s->prims[i].ntris : 1;
    if (total == 0) return -1;
    comp_info *ci = (comp_info *)malloc(total * sizeof(comp_info));
    size_t k = 0;
    for (uint32_t i = 0; i < s->num_prims; ++i) {
        uint32_t nc = s->prims[i].type == ORC_PRIM_MESH ? s->prims[i].ntris : 1;
        for (uint32_t c = 0; c < nc; ++c) {
            ci[k].prim = i;
            ci[k].comp = c;
            ci[k].bbox = prim_bbox(&s->prims[i], c);
            ci[k].centroid = v_add(v_scale(ci[k].bbox.min, 0.5f), v_scale(ci[k].bbox.max, 0.5f));
            ++k;
        }
    }
    build_ctx bc = {0};
    bc.ordered = (uint32_t *)malloc(2 * total * sizeof(uint32_t));
    uint32_t root = recurse_build(&bc, ci, total);
    s->nodes = (orc_node *)malloc(bc.n * sizeof(orc_node));
    uint32_t count = 0;
    flatten(&bc, s->nodes, &count, root);
    s->num_nodes = count;
    s->comps = bc.ordered;
    s->num_comps = (uint32_t)bc.n_ordered;
    free(bc.nodes);
    free(ci);
    s->lights = (uint32_t *)malloc(sizeof(uint32_t) * (s->num_prims + 1));
    s->num_lights = 0;
    for (uint32_t i = 0; i < s->num_prims; ++i)
        if (s->prims[i].mat.has_light) s->lights[s->num_lights++] = i;
    s->slab_pad = slab_pad();
    return 0;
}

uint32_t orc_bvh_num_nodes(const orc_scene *s) { return s->num_nodes; }
uint32_t orc_bvh_num_components(const orc_scene *s) { return s->num_comps; }
uint32_t orc_num_lights(const orc_scene *s) { return s->num_lights; }
void orc_light_indices(const orc_scene *s, uint32_t *out)
{
    for (uint32_t i = 0; i < s->num_lights; ++i) out[i] = s->lights[i];
}

void orc_bvh_export(const orc_scene *s, float *bbox6, uint32_t *offset, uint32_t *ncomp,
                    uint32_t *axis, uint32_t *comp_prim, uint32_t *comp_index)
{
    for (uint32_t i = 0; i < s->num_nodes; ++i) {
        const orc_node *nd = &s->nodes[i];
        bbox6[6*i+0] = nd->bbox.min.x; bbox6[6*i+1] = nd->bbox.min.y; bbox6[6*i+2] = nd->bbox.min.z;
        bbox6[6*i+3] = nd->bbox.max.x; bbox6[6*i+4] = nd->bbox.max.y; bbox6[6*i+5] = nd->bbox.max.z;
        offset[i] = nd->offset; ncomp[i] = nd->n; axis[i] = nd->axis;
    }
    for (uint32_t i = 0; i < s->num_comps; ++i) {
        comp_prim[i] = s->comps[2*i];
        comp_index[i] = s->comps[2*i+1];
    }
}

/* ------------------------------------------------------------------------ */
/* Traversal                                                                 */
/* ------------------------------------------------------------------------ */

static float comp_intersect(const orc_scene *s, uint32_t prim, uint32_t comp, ovec o, ovec d,
                            uint32_t *cnt_tri, uint32_t *cnt_sph)
{
    const orc_prim *p = &s->prims[prim];
    if (p->type == ORC_PRIM_MESH) { (*cnt_tri)++; return orc_tri_intersect(p, comp, o, d, NULL, NULL); }
    (*cnt_sph)++;
    return orc_sphere_intersect(p, o, d);
}

#define MAX_NODES_TO_VISIT 64

/* Bvh::intersect, src/geom/bvh.rs:406-456.  Returns 0, or -1 when the 64-entry
 * VisitStack would overflow (the reference panics, bvh.rs:156-157).
 * stop_below: the traversal ends at the first accepted hit closer than this
 * (the early exit of a visibility query that only needs one occluder closer
 * than its target; -INFINITY never stops early, as the reference). */
static int intersect_impl(const orc_scene *s, ovec o, ovec d, orc_hit *hit, uint32_t cnt[3],
                          float stop_below)
{
    float closest = FLT_MAX;
    hit->prim = ORC_NONE; hit->comp = ORC_NONE; hit->dist = 0.0f;
    /* Ray::compute_intersection_data, src/core/ray.rs:31-42. */
    ovec inv = v_new(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int neg[3] = {d.x < 0.0f, d.y < 0.0f, d.z < 0.0f};
    uint32_t stack[MAX_NODES_TO_VISIT];
    uint32_t sp = 0, cur = 0;
    uint32_t c_box = 0, c_tri = 0, c_sph = 0;
    for (;;) {
        const orc_node *nd = &s->nodes[cur];
        c_box++;
        if (bb_intersect(&nd->bbox, o, inv, neg, closest, s->slab_pad)) {
            if (nd->n > 0) {
                for (uint32_t i = nd->offset; i < nd->offset + nd->n; ++i) {
                    uint32_t pi = s->comps[2*i], ci = s->comps[2*i+1];
                    float dist = comp_intersect(s, pi, ci, o, d, &c_tri, &c_sph);
                    if (dist != 0.0f && dist < closest) {
                        hit->prim = pi; hit->comp = ci; hit->dist = dist;
                        closest = dist;
                        if (dist < stop_below) goto done;
                    }
                }
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                if (sp == MAX_NODES_TO_VISIT) {
                    hit->prim = ORC_NONE; hit->comp = ORC_NONE; hit->dist = 0.0f;
                    return -1;
                }
                if (neg[nd->axis]) { stack[sp++] = cur + 1; cur = nd->offset; }
                else { stack[sp++] = nd->offset; cur = cur + 1; }
            }
        } else {
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
done:
    if (cnt) { cnt[0] = c_box; cnt[1] = c_tri; cnt[2] = c_sph; }
    return 0;
}

int orc_intersect_one(const orc_scene *s, ovec o, ovec d, orc_hit *hit, uint32_t cnt[3])
{
    return intersect_impl(s, o, d, hit, cnt, -INFINITY);
}

/* Bvh::intersect_naive, src/geom/bvh.rs:385-400. */
static void intersect_naive_one(const orc_scene *s, ovec o, ovec d, orc_hit *hit)
{
    float closest = FLT_MAX;
    uint32_t dummy1 = 0, dummy2 = 0;
    hit->prim = ORC_NONE; hit->comp = ORC_NONE; hit->dist = 0.0f;
    for (uint32_t pi = 0; pi < s->num_prims; ++pi) {
        uint32_t nc = s->prims[pi].type == ORC_PRIM_MESH ? s->prims[pi].ntris : 1;
        for (uint32_t c = 0; c < nc; ++c) {
            float dist = comp_intersect(s, pi, c, o, d, &dummy1, &dummy2);
            if (dist != 0.0f && dist < closest) {
                hit->prim = pi; hit->comp = c; hit->dist = dist;
                closest = dist;
            }
        }
    }
}

int64_t orc_intersect(const orc_scene *s, const float *rays, int64_t n, uint32_t *prim,
                      uint32_t *comp, float *dist, uint32_t *counts, float *frame, int nthreads)
{
    int64_t overflow = 0;
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256) reduction(+:overflow)
    for (int64_t i = 0; i < n; ++i) {
        const float *r = &rays[6 * i];
        ovec o = v_new(r[0], r[1], r[2]), d = v_new(r[3], r[4], r[5]);
        orc_hit h;
        uint32_t cnt[3] = {0, 0, 0};
        int ovf = orc_intersect_one(s, o, d, &h, cnt) != 0;
        if (ovf) { overflow++; h.dist = -1.0f; }
        prim[i] = h.prim; comp[i] = h.comp; dist[i] = h.dist;
        if (counts) { counts[3*i] = cnt[0]; counts[3*i+1] = cnt[1]; counts[3*i+2] = cnt[2]; }
        if (frame) {
            orc_frame f;
            if (h.prim != ORC_NONE) orc_hit_frame(s, &h, o, d, &f);
            else memset(&f, 0, sizeof(f));
            float *fo = &frame[12 * i];
            fo[0] = f.normal.x; fo[1] = f.normal.y; fo[2] = f.normal.z;
            fo[3] = f.tangent.x; fo[4] = f.tangent.y; fo[5] = f.tangent.z;
            fo[6] = f.binormal.x; fo[7] = f.binormal.y; fo[8] = f.binormal.z;
            fo[9] = f.geom_normal.x; fo[10] = f.geom_normal.y; fo[11] = f.geom_normal.z;
        }
    }
    return overflow;
}

void orc_intersect_naive(const orc_scene *s, const float *rays, int64_t n, uint32_t *prim,
                         uint32_t *comp, float *dist, int nthreads)
{
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float *r = &rays[6 * i];
        orc_hit h;
        intersect_naive_one(s, v_new(r[0], r[1], r[2]), v_new(r[3], r[4], r[5]), &h);
        prim[i] = h.prim; comp[i] = h.comp; dist[i] = h.dist;
    }
}

/* Bvh::visibility, src/geom/bvh.rs:460-477.  Returns 1 visible, 0 occluded,
 * -1 on stack overflow.  early_exit = 0 runs the reference's full closest-hit
 * traversal; 1 stops at the first accepted hit closer than target_dist - 1e-3
 * (same answer: once the closest distance is below that bound it stays below
 * it), so cnt then holds the work of a traversal that ends at its first
 * occluder. */
static int visibility_impl(const orc_scene *s, ovec start, ovec target, uint32_t cnt[3], int early_exit)
{
    if (cnt) { cnt[0] = cnt[1] = cnt[2] = 0; }
    if (v_is_close(start, target, 1e-3f)) return 0;
    ovec dir = v_normalized(v_sub(target, start));
    ovec origin = v_add(start, v_scale(dir, ORC_RAY_PUSH_DIST)); /* Ray::nudge, ray.rs:44-46 */
    float target_dist = v_mag(v_sub(target, origin));
    orc_hit h;
    if (intersect_impl(s, origin, dir, &h, cnt, early_exit ? target_dist - 1e-3f : -INFINITY) != 0) return -1;
    if (h.prim != ORC_NONE) {
        if (h.dist < (target_dist - 1e-3f)) return 0;
    }
    return 1;
}

int orc_visibility_one(const orc_scene *s, ovec start, ovec target, uint32_t cnt[3])
{
    return visibility_impl(s, start, target, cnt, 0);
}

int64_t orc_visibility(const orc_scene *s, const float *segs, int64_t n, uint8_t *out,
                       uint32_t *counts, int early_exit, int nthreads)
{
    int64_t overflow = 0;
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256) reduction(+:overflow)
    for (int64_t i = 0; i < n; ++i) {
        const float *g = &segs[6 * i];
        uint32_t cnt[3];
        int v = visibility_impl(s, v_new(g[0], g[1], g[2]), v_new(g[3], g[4], g[5]), cnt, early_exit);
        if (v < 0) { overflow++; v = 0; }
        out[i] = (uint8_t)v;
        if (counts) { counts[3*i] = cnt[0]; counts[3*i+1] = cnt[1]; counts[3*i+2] = cnt[2]; }
    }
    return overflow;
}
