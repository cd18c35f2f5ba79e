/*
 * gr_oracle.h -- CPU oracle for the gammaray ray-traversal + integrator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * renderer (sdao/gammaray, Rust) that the GPU library is checked against.  Only
 * the test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it, and only as the checker / the timed CPU baseline.  The product
 * library never links, loads or calls anything in this directory.
 *
 * Every function cites the reference file:line it restates.  The reference is
 * f32 throughout and Rust never contracts a*b+c into an FMA, so this file is
 * compiled with -ffp-contract=off -fno-fast-math and keeps the reference's
 * evaluation order.  f32::min/max map to fminf/fmaxf (same NaN rule).
 *
 * Parity pins (what checks each part; see tests/):
 *   - BVH build/flatten: structural invariants + hand-derived small trees.
 *   - intersect: the reference's own brute force (intersect_naive) restated,
 *     closed-form ray/triangle and ray/sphere hits, axis-aligned edge cases.
 *   - visibility: equivalence with a brute-force segment test.
 *   - RNG: Philox4x32-10 published known-answer vectors.
 *   - samplers / lobes / integrators: closed forms, normalisation, chi-square
 *     and analytic scenes (listed in the test files).
 */
#ifndef GR_ORACLE_H
#define GR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Material record: 16 floats.  rec[0] = kind:
 *   0 = Disney (rec[1..3] base colour, [4] roughness, [5] anisotropic, [6] ior,
 *       [7] metallic, [8] specular_trans, [9] specular_tint, [10] sheen,
 *       [11] sheen_tint, [12] clearcoat, [13] clearcoat_gloss)
 *   1 = diffuse_light (rec[1..3] incandescence)
 *   2 = diffuse (PerfectDiffuse)      3 = mirror (PerfectMirror)          */
#define ORC_MAT_FLOATS 16

int  orc_scene_create(orc_scene **out);
void orc_scene_destroy(orc_scene *s);
/* Mesh from arrays (the reference only loads OBJ; this is the array form of
 * Mesh::from_obj's result, src/geom/mesh.rs:93-188).  tris holds 9 uint32 per
 * triangle: a,b,c (vertex), an,bn,cn (normal; 0xffffffff = use geometric
 * normal), at,bt,ct (uv; 0xffffffff = zero uv).  Returns the prim index. */
int  orc_scene_add_mesh(orc_scene *s, const float *verts, uint32_t nverts,
                        const float *normals, uint32_t nnormals,
                        const float *uvs, uint32_t nuvs,
                        const uint32_t *tris, uint32_t ntris, const float *mat);
int  orc_scene_add_sphere(orc_scene *s, const float *center, float radius, const float *mat);
int  orc_scene_build(orc_scene *s);

/* Flattened tree export (BvhLinearNode, src/geom/bvh.rs:87-92). */
uint32_t orc_bvh_num_nodes(const orc_scene *s);
uint32_t orc_bvh_num_components(const orc_scene *s);
void orc_bvh_export(const orc_scene *s, float *bbox6, uint32_t *offset, uint32_t *ncomp,
                    uint32_t *axis, uint32_t *comp_prim, uint32_t *comp_index);
uint32_t orc_num_lights(const orc_scene *s);
void orc_light_indices(const orc_scene *s, uint32_t *out);

/* Batched closest hit (Bvh::intersect, src/geom/bvh.rs:406-456).  rays: 6 floats
 * (origin, direction).  Outputs per ray: prim (0xffffffff = miss), component,
 * dist.  counts (optional, 3 per ray): bbox tests, triangle tests, sphere tests.
 * frame (optional, 12 per ray): normal, tangent, binormal, geom_normal.
 * Returns the number of rays whose traversal overflowed the 64-entry stack
 * (the reference panics there); those rays report prim = 0xffffffff and
 * dist = -1. */
int64_t orc_intersect(const orc_scene *s, const float *rays, int64_t n, uint32_t *prim,
                      uint32_t *comp, float *dist, uint32_t *counts, float *frame, int nthreads);
/* Brute force (Bvh::intersect_naive, src/geom/bvh.rs:385-400). */
void orc_intersect_naive(const orc_scene *s, const float *rays, int64_t n, uint32_t *prim,
                         uint32_t *comp, float *dist, int nthreads);
/* Bvh::visibility (src/geom/bvh.rs:460-477).  segs: 6 floats (start, target).
 * out: 1 = visible.  counts optional as for orc_intersect.  early_exit = 0:
 * the reference's full closest-hit traversal; 1: the traversal ends at the
 * first accepted hit closer than the target (same answers; counts then give
 * the work of a first-occluder traversal). */
int64_t orc_visibility(const orc_scene *s, const float *segs, int64_t n, uint8_t *out,
                       uint32_t *counts, int early_exit, int nthreads);

/* ---- rendering (orc_render.c) ---- */
#define ORC_PT 0      /* PathTracerIntegrator + max_depth cap */
#define ORC_PT_NEE 1  /* path tracing with next-event estimation + MIS */
#define ORC_BDPT 2    /* BdptIntegrator with camera/light vertex caps */
#define ORC_DISPLAY 4 /* DisplayColorIntegrator, integrators.rs:18-34 */
#define ORC_BDPT_MIS 5 /* BDPT over the same strategies with power-heuristic MIS weights (no reference counterpart) */
#define ORC_BDPT_LT 6  /* ORC_BDPT_MIS plus the light-tracing strategy (light vertices splatted onto the film) */
/* Camera::new's xform = translation * rotation (row-major 4x4). */
void orc_camera_matrix(const float quat[4], const float translate[3], float out[16]);
/* Camera::compute_ray for n lens points st (2 floats each) -> rays (6 each). */
void orc_camera_rays(const float cam[16], float wx, float wy, const float *st, int64_t n, float *rays);
float orc_mitchell1(float x);
/* Film::compute_sample_points for iteration `sample_offset` (2 floats/pixel). */
void orc_sample_points(uint64_t seed, uint64_t sample_offset, uint32_t width, uint32_t height, float *st);
/* Film::report_samples for n samples (st, colors) in order: film holds
 * (r, g, b, weight) per pixel. */
void orc_report_samples(float *film, uint32_t width, uint32_t height, const float *st, const float *colors,
                        int64_t n);
/* Expected per-iteration film weight of each of n pixel columns (or rows). */
void orc_expected_weights(uint32_t n, float *out);
/* spp iterations of Stage::trace: per iteration, sample points, integrate
 * every pixel's sample (path id = iteration * W * H + pixel), splat in sample
 * order.  colors (optional) receives spp * W * H * 3 sample radiances. */
int orc_render(const orc_scene *s, const float cam[16], float wx, float wy, uint32_t width,
               uint32_t height, int integrator, uint32_t max_depth, uint32_t light_depth,
               uint64_t seed, uint64_t sample_offset, uint32_t spp, float *film, float *colors,
               int nthreads);
void orc_material_query(const orc_scene *s, uint32_t prim, int64_t n, const float *in, float *out);
/* ExrWriter bytes for a film (rgbw per pixel); returns the size, writes up to
 * cap bytes to out (may be NULL). */
size_t orc_exr_encode(const float *rgbw, uint32_t width, uint32_t height, uint8_t *out, size_t cap);

/* Philox4x32-10 block function (counter-based RNG shared by the GPU path). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* BBox::intersect per (box, ray, max_dist); boxes lo.xyz hi.xyz, rays o.xyz d.xyz. */
void orc_bbox_intersect(const float *boxes, const float *rays, const float *max_dist, int64_t n,
                        uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
