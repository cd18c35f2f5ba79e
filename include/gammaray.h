/*
 * gammaray.h -- C ABI of the MI355X-native gammaray hot path.
 *
 * The reference (sdao/gammaray, Rust) has no FFI; its per-ray, per-sample
 * interfaces are too fine-grained for a GPU.  This ABI sits at the granularity
 * a Rust (or any) host would bind through `extern "C"`:
 *   - scene construction  <->  Stage::new(Vec<Box<Prim>>)          src/render/stage.rs:16-21
 *   - batched traversal   <->  Bvh::intersect / Bvh::visibility    src/geom/bvh.rs:406-477
 *   - one render pass     <->  Stage::trace(&Camera, &Integrator, &mut Film)
 *                                                                  src/render/stage.rs:23-36
 *   - film readback       <->  Film::pixels / ExrWriter::update    src/render/film.rs:38-115,
 *                                                                  src/render/exr.rs:168-199
 * Plain pointers and sizes only.  Every entry point returns GR_OK (0) or a
 * negative GR_ERR_* code and never aborts; gr_ctx_last_error() gives the text.
 * A context owns one HIP stream; calls on one context are stream-ordered and a
 * context must be used by one host thread at a time.
 */
#ifndef GAMMARAY_H
#define GAMMARAY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GR_OK 0
#define GR_ERR_INVALID (-1)   /* bad argument */
#define GR_ERR_HIP (-2)       /* HIP runtime failure */
#define GR_ERR_NOMEM (-3)     /* allocation failure */
#define GR_ERR_STATE (-4)     /* call out of order (e.g. scene not committed) */
#define GR_ERR_OVERFLOW (-5)  /* a traversal stack overflowed (reference: panic, bvh.rs:156) */

#define GR_NONE 0xffffffffu

typedef struct gr_ctx gr_ctx;
typedef struct gr_scene gr_scene;
typedef struct gr_film gr_film;

/* Ray: origin, pad, unit direction, pad (32 bytes; src/core/ray.rs:11-14). */
typedef struct { float ox, oy, oz, pad0, dx, dy, dz, pad1; } gr_ray;
/* Closest hit.  prim/comp = (prim index, component index) or GR_NONE on a miss;
 * t = distance along the ray (Intersection::Hit.dist, src/geom/bvh.rs:108-115
 * plus the component index the reference drops at bvh.rs:422). */
typedef struct { uint32_t prim, comp; float t; uint32_t slot; } gr_hit;
/* Shadow segment: start, pad, target, pad (arguments of Bvh::visibility). */
typedef struct { float sx, sy, sz, pad0, tx, ty, tz, pad1; } gr_segment;

/* ---- context ---------------------------------------------------------- */
/* device >= 0: a HIP device.  device == GR_HOST_ONLY: no GPU is touched; scenes
 * can be built and their tree exported, but nothing can be traced. */
#define GR_HOST_ONLY (-1)
int gr_ctx_create(int device, gr_ctx **out);
/* Releases the caller's handle.  Scenes, films and guides created on the
 * context keep it alive: it is freed when the last of them is destroyed, so
 * they may be destroyed after the context in any order. */
void gr_ctx_destroy(gr_ctx *ctx);
/* Run subsequent work on an existing HIP stream (hipStream_t; NULL is the
 * legacy default stream).  gr_ctx_use_own_stream() returns to the context's
 * own (blocking) stream. */
int gr_ctx_set_stream(gr_ctx *ctx, void *hip_stream);
int gr_ctx_use_own_stream(gr_ctx *ctx);
void *gr_ctx_stream(gr_ctx *ctx);
int gr_ctx_synchronize(gr_ctx *ctx);
const char *gr_ctx_last_error(gr_ctx *ctx);

/* Device-side error counters of the context's render calls (gr_trace,
 * gr_render_jobs), accumulated since the previous gr_ctx_take_stats.  They
 * stand in for the reference's debug assertions; a release build of the
 * reference lets the values through, and so does this library (a non-finite
 * sample reaches the film as in the reference). */
typedef struct {
    uint64_t nonfinite_bsdf_samples; /* material samples whose radiance or pdf is not finite, or whose
                                        pdf is negative (debug_assert!s, src/material/material.rs:224-226) */
    uint64_t nonfinite_radiance;     /* camera samples whose radiance is not finite when they reach the
                                        film (Film::report_samples, src/render/film.rs:78-114) */
    uint64_t bad_splats;             /* light-tracing contributions not finite (dropped) or beyond the
                                        fixed-point splat range (clamped); GR_INTEGRATOR_BDPT_LT only */
    uint64_t reserved;
} gr_stats;
/* Waits for the context's stream, returns the counters and clears them. */
int gr_ctx_take_stats(gr_ctx *ctx, gr_stats *out);

/* ---- scene (Stage::new / Bvh::build) ------------------------------------ */
/* Material record: 16 floats, rec[0] = kind:
 *   0 Disney: [1..3] base colour, [4] roughness, [5] anisotropic, [6] ior,
 *     [7] metallic, [8] specular_trans, [9] specular_tint, [10] sheen,
 *     [11] sheen_tint, [12] clearcoat, [13] clearcoat_gloss
 *     (DisneyMaterialBuilder, src/material/material.rs:256-389)
 *   1 diffuse_light: [1..3] incandescence (material.rs:27-33)
 *   2 diffuse (material.rs:35-43)       3 mirror (material.rs:45-53)        */
#define GR_MAT_FLOATS 16

int gr_scene_create(gr_ctx *ctx, gr_scene **out);
void gr_scene_destroy(gr_scene *scene);
/* Triangle mesh in world space (the state Mesh::from_obj builds,
 * src/geom/mesh.rs:93-188).  tris: 9 uint32 per triangle (a,b,c vertex;
 * an,bn,cn normal or GR_NONE = geometric normal; at,bt,ct uv or GR_NONE = zero
 * uv).  Writes the prim index to *prim_out (may be NULL). */
int gr_scene_add_mesh(gr_scene *scene, const float *verts, uint32_t nverts,
                      const float *normals, uint32_t nnormals, const float *uvs, uint32_t nuvs,
                      const uint32_t *tris, uint32_t ntris, const float *material,
                      uint32_t *prim_out);
/* Sphere::new(material, translation, radius), src/geom/sphere.rs:17-24. */
int gr_scene_add_sphere(gr_scene *scene, const float center[3], float radius,
                        const float *material, uint32_t *prim_out);
/* Wavefront OBJ, parsed on the host into the state Mesh::from_obj builds
 * (src/geom/mesh.rs:61-188): vertices through Xform::transform, normals
 * through Xform::transform_normal (Mat::inverted) and normalised unless nearly
 * zero, n-gons as fans (0, k, k+1).  xf: 16 floats, the row-major Mat
 * (xf[4 r + c] = mat[r][c]); NULL = identity.  On failure err (if given)
 * receives the message, e.g. "OBJ parse error at line N: ...". */
typedef struct gr_obj gr_obj;
int gr_obj_load(const char *path, const float *xf, gr_obj **out, char *err, size_t err_len);
int gr_obj_sizes(const gr_obj *obj, uint32_t *nverts, uint32_t *nnormals, uint32_t *nuvs,
                 uint32_t *ntris);
/* verts: 3 floats each, normals 3, uvs 2, tris 9 uint32 each (a,b,c vertex;
 * an,bn,cn normal or GR_NONE = geometric normal; at,bt,ct uv or GR_NONE = zero
 * uv), the layout gr_scene_add_mesh takes.  Any output may be NULL. */
int gr_obj_read(const gr_obj *obj, float *verts, float *normals, float *uvs, uint32_t *tris);
void gr_obj_destroy(gr_obj *obj);
/* gr_scene_add_mesh with the loaded arrays (Mesh::from_obj + Stage::new). */
int gr_scene_add_obj(gr_scene *scene, const gr_obj *obj, const float *material, uint32_t *prim_out);
/* Bvh::build (src/geom/bvh.rs:339-382) on the host -- the reference's SAH
 * builder, so the tree is identical -- then upload of the GPU layout. */
int gr_scene_commit(gr_scene *scene);

typedef struct {
    uint32_t num_prims, num_components, num_nodes, num_interior, num_lights, max_depth;
    uint64_t node_bytes, component_bytes;
    double build_seconds;
} gr_scene_info;
int gr_scene_info_get(gr_scene *scene, gr_scene_info *out);
/* Copy of the flattened tree in the reference's BvhLinearNode order
 * (src/geom/bvh.rs:87-106,310-337): bbox6 (6 floats per node), offset, count,
 * axis per node; (prim, comp) per ordered component.  Any pointer may be NULL. */
int gr_scene_export_bvh(gr_scene *scene, float *bbox6, uint32_t *offset, uint32_t *count,
                        uint32_t *axis, uint32_t *comp_prim, uint32_t *comp_index);

/* ---- batched traversal (device pointers) -------------------------------- */
/* Bvh::intersect for n rays.  d_rays, d_hits: device memory.  If d_counts is
 * not NULL it receives 3 uint32 per ray: bbox tests, triangle tests, sphere
 * tests (the reference-order traversal's work, for the roofline's bytes). */
int gr_intersect_closest(gr_scene *scene, const gr_ray *d_rays, uint32_t n, gr_hit *d_hits,
                         uint32_t *d_counts);
/* gr_intersect_closest without any allocation.  No counterpart in the
 * reference (it has no allocator to avoid); for callers that cannot allocate:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments and results as gr_intersect_closest, with the
 * scene's tuning (ray sort included).  The ray-sort scratch and the traversal's
 * work counters live in one static device array per GPU (about 16.3 MiB, held
 * until the process ends), so n is limited to GR_NOALLOC_MAX_RAYS (larger n:
 * GR_ERR_INVALID).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams or host threads) race. */
#define GR_NOALLOC_MAX_RAYS (1u << 20)
int gr_intersect_closest_noalloc(gr_scene *scene, const gr_ray *d_rays, uint32_t n, gr_hit *d_hits,
                                 uint32_t *d_counts);
/* Bvh::visibility for n segments: d_visible[i] = 1 if target is visible from
 * start, else 0.  mode GR_OCCLUSION_EXACT runs the reference's closest-hit
 * traversal order and stops at the first hit that proves occlusion (result
 * identical to the reference on every input); GR_OCCLUSION_ANYHIT also prunes
 * nodes beyond the target (identical except documented ties within rounding of
 * the target distance). */
#define GR_OCCLUSION_EXACT 0
#define GR_OCCLUSION_ANYHIT 1
int gr_occluded(gr_scene *scene, const gr_segment *d_segs, uint32_t n, uint32_t *d_visible,
                int mode, uint32_t *d_counts);
/* Number of rays whose traversal stack overflowed since the last call (the
 * reference panics on these, src/geom/bvh.rs:156-157).  Such rays report
 * prim = GR_NONE with t = -1 (closest hit) or not-visible (visibility). */
int gr_scene_take_overflow_count(gr_scene *scene, uint64_t *out);

/* ---- traversal tuning (no reference counterpart: the reference traverses one
 * ray at a time on the CPU, src/geom/bvh.rs:406-477) ------------------------
 * Applies to every traversal of the scene: the batched calls above and the
 * rays gr_trace casts.  No setting changes a result, only speed. */
typedef struct {
    uint32_t sort_bits;    /* 0: rays run in input order; else rays are sorted by a key of
                              this many bits (6..30, rounded down to 3 + 3m): direction
                              octant above the m-bit-per-axis Morton code of the origin */
    uint32_t lds_stack;    /* traversal stack entries per lane kept in LDS: 8 (default), 16 or 32 */
    uint32_t waves_per_cu; /* resident traversal waves per CU; 0 = occupancy limit */
    uint32_t refill;       /* idle lanes (1..64) at which a closest-hit wave fetches new rays
                              (default 32) */
    uint32_t xcd_ranges;   /* 1 (default): the (sorted) batch is cut into one contiguous range
                              per XCD, each drained by the waves resident on that XCD first;
                              0: one range */
    uint32_t ray_chunk;    /* rays a wave reserves per atomic on the work counter (1..65536);
                              0 = 128 */
    uint32_t vis_refill;   /* idle lanes (1..64) at which a visibility wave fetches new
                              segments (default 40); 0 = refill */
} gr_tuning;
/* A new scene's ray batches (gr_intersect_*, gr_visibility) run the defaults
 * above; its integrators (gr_trace, gr_render_jobs without a tuning) run
 * refill 16, vis_refill 16, ray_chunk 64, which suit their shrinking bounce
 * queues.  gr_scene_set_tuning and gr_scene_autotune set the one setting every
 * launch on the scene then runs; gr_scene_get_tuning reports the ray batches'. */
int gr_scene_set_tuning(gr_scene *scene, const gr_tuning *tuning);
int gr_scene_get_tuning(gr_scene *scene, gr_tuning *out);
/* Times candidate settings on n rays (device memory; closest hit) and keeps the
 * fastest: sort bits (12, 15, 18, 21, 24, 30), XCD ranges, LDS stack depth,
 * waves per CU, refill, ray chunk, visibility refill.
 * d_hits: scratch for n hits.  The chosen setting is returned in *out. */
int gr_scene_autotune(gr_scene *scene, const gr_ray *d_rays, uint32_t n, gr_hit *d_hits,
                      gr_tuning *out);
/* The same search on a mixed workload: each candidate is timed on n closest-hit
 * rays plus m visibility segments (mode as gr_occluded), since one setting
 * serves both kinds of launch.  d_visible: scratch for m results.  n or m may
 * be 0 (not both). */
int gr_scene_autotune_mixed(gr_scene *scene, const gr_ray *d_rays, uint32_t n, gr_hit *d_hits,
                            const gr_segment *d_segs, uint32_t m, uint32_t *d_visible, int mode,
                            gr_tuning *out);
/* The same search for gr_render_jobs (no reference counterpart): candidates
 * are timed with the batch's multi-scene traversal on n rays and m visibility
 * segments (exact mode) over the nscenes scenes, each ray and segment naming
 * its scene (an index into `scenes`) in the pad word after its origin, as the
 * batch's own do.  The setting is returned in *out (pass it as the batch's
 * tuning); no scene's own setting changes. */
int gr_jobs_autotune(gr_ctx *ctx, gr_scene *const *scenes, uint32_t nscenes, const gr_ray *d_rays, uint32_t n,
                     gr_hit *d_hits, const gr_segment *d_segs, uint32_t m, uint32_t *d_visible,
                     gr_tuning *out);

/* ---- kernel timing (observability; no reference counterpart) -------------
 * When enabled, every traversal and ray-sort launch on the context is
 * bracketed by HIP events on the stream it runs on (do not enable while a
 * caller captures the stream into a graph). */
#define GR_KERNEL_TRACE_CLOSEST 0
#define GR_KERNEL_TRACE_SHADOW 1
#define GR_KERNEL_RAY_SORT 2
#define GR_KERNEL_KINDS 3
int gr_ctx_set_kernel_timing(gr_ctx *ctx, int enable);
/* Waits for the context's stream; then, per kind, the summed milliseconds and
 * the number of launches since the previous call (either pointer may be NULL). */
int gr_ctx_take_kernel_times(gr_ctx *ctx, double *ms, uint64_t *launches);

/* ---- rendering (Stage::trace, src/render/stage.rs:23-36) ---------------- */
/* Camera::new arguments (src/core/camera.rs:42-61); f_stop is carried but, as
 * in the reference, unused (pinhole camera). */
typedef struct {
    float focal_length, horizontal_aperture, vertical_aperture, f_stop;
    float rotate[4];    /* quaternion (real, i, j, k) */
    float translate[3];
} gr_camera;

#define GR_INTEGRATOR_PT 0      /* PathTracerIntegrator, src/render/integrators.rs:39-100 */
#define GR_INTEGRATOR_PT_NEE 1  /* path tracing + next-event estimation with MIS (no reference counterpart) */
#define GR_INTEGRATOR_BDPT 2    /* BdptIntegrator, src/render/integrators.rs:102-418 */
#define GR_INTEGRATOR_DISPLAY 4 /* DisplayColorIntegrator, src/render/integrators.rs:18-34: the display
                                   colour of the first hit's material, black on a miss */
#define GR_INTEGRATOR_GUIDED 3  /* PT_NEE whose BSDF sampling is mixed (one-sample MIS) with a
                                   learned guide trained online (no reference counterpart) */
#define GR_INTEGRATOR_BDPT_MIS 5 /* BdptIntegrator's subpaths and connections (integrators.rs:102-418) with
                                    power-heuristic MIS weights over the strategies in place of the
                                    uniform 1/(c+l) of BdptIntegrator::weight (integrators.rs:345-354) */
#define GR_INTEGRATOR_BDPT_LT 6  /* GR_INTEGRATOR_BDPT_MIS plus the light-tracing strategy: every light
                                    vertex the camera sees is joined to it and splatted onto the film with
                                    the film's filter; the power-heuristic weights include that strategy.
                                    d_sample_radiance then holds the camera strategies' part only */

typedef struct gr_guide gr_guide;

typedef struct {
    uint32_t integrator;
    uint32_t spp;            /* iterations; one sample per pixel each (film.rs:55-76) */
    uint32_t max_depth;      /* intersections per camera path; 0 = unbounded (the reference) */
    uint32_t light_depth;    /* BDPT: bounces after the light vertex; 0 = reference cap */
    uint64_t seed;           /* RNG key */
    uint64_t sample_offset;  /* index of the first iteration (progressive rendering, resume) */
    uint32_t batch;          /* iterations traced together; 0 = automatic */
    uint32_t reserved;
    float *d_sample_radiance; /* optional device buffer: spp*W*H float4 per-sample radiance */
    gr_guide *guide;         /* GR_INTEGRATOR_GUIDED: the guide to sample and train */
} gr_trace_params;

/* FilmPixel grid (src/render/film.rs:27-53): per pixel (accum rgb, weight). */
int gr_film_create(gr_ctx *ctx, uint32_t width, uint32_t height, gr_film **out);
void gr_film_destroy(gr_film *film);
int gr_film_clear(gr_film *film);
/* Host copy of the film: width*height*4 floats (r, g, b, weight), row 0 first. */
int gr_film_readback(gr_film *film, float *rgbw);
/* Overlapped readback (the progressive loop of src/bin/gammaray.rs:85-112):
 * begin snapshots the film as it will be after the work queued so far and
 * starts its copy to host memory on a side stream, so later gr_trace calls
 * render meanwhile; end waits for that copy and writes it to rgbw
 * (width*height*4 floats).  One readback in flight per film. */
int gr_film_readback_begin(gr_film *film);
int gr_film_readback_end(gr_film *film, float *rgbw);
/* ExrWriter::update + write from a host film copy (no GPU needed). */
int gr_exr_write(const float *rgbw, uint32_t width, uint32_t height, const char *path);
/* Restores a film from a host copy (checkpoint / resume). */
int gr_film_upload(gr_film *film, const float *rgbw);
/* spp iterations of Stage::trace with the given integrator; samples splat into
 * the film with the reference's Mitchell filter in the reference's order. */
int gr_trace(gr_scene *scene, const gr_camera *camera, const gr_trace_params *params, gr_film *film);
/* ExrWriter::update + write (src/render/exr.rs:168-199): scanline EXR, B,G,R
 * float channels, rows flipped, value = accum / weight. */
int gr_film_write_exr(gr_film *film, const char *path);

/* ---- batched rendering (no reference counterpart: the reference renders one
 * job, src/bin/gammaray.rs:85-112) -------------------------------------------
 * Many jobs -- each a scene, camera, film, spp and depth cap -- run through one
 * shared pool of in-flight paths.  One bounce step serves every job at once:
 * the room the continuing paths leave in the queue takes the next samples
 * from a global counter over all jobs, one closest-hit launch traverses every
 * job's scene (each ray names its scene), shading and shadow rays follow,
 * finished paths hand their radiance to their job.  The step is captured once as a hipGraph
 * and replayed.  Finished samples wait in a ring of sample_ring entries (24 bytes each); each
 * job's film gets the Mitchell splat of its iterations, in order, as soon as an iteration's
 * samples are all in, so memory is bounded by the pool and the ring, not by the jobs' total
 * samples.  A job's film equals what gr_trace renders for the same camera, seed, spp, depth
 * and sample_offset. */
typedef struct {
    gr_scene *scene;
    gr_film *film;             /* the resolution is the film's */
    gr_camera camera;
    uint32_t spp, max_depth;   /* max_depth 0 = unbounded, as gr_trace */
    uint64_t seed, sample_offset;
} gr_job;
typedef struct {
    uint32_t integrator;       /* GR_INTEGRATOR_PT or GR_INTEGRATOR_PT_NEE, for every job */
    uint32_t path_pool;        /* paths in flight; 0 = 8,388,608 */
    uint32_t steps_per_poll;   /* bounce steps between completion checks; 0 = 16 */
    uint32_t use_graph;        /* 1: replay the captured bounce step as a hipGraph; 0: launch eagerly */
    gr_tuning tuning;          /* the shared traversal launches; lds_stack 0 = the first job's scene's setting */
    uint32_t job_order;        /* order in which jobs take samples: GR_JOBS_AS_GIVEN or GR_JOBS_FEWEST_FIRST */
    uint32_t sample_ring;      /* finished samples held before their film gather; 0 = 16 x path_pool
                                  (at most the batch's samples); raised to the largest job's pixel
                                  count if smaller */
} gr_batch_params;
#define GR_JOBS_AS_GIVEN 0
#define GR_JOBS_FEWEST_FIRST 1 /* fewest samples first (shorter jobs finish sooner: lower median latency) */
/* All scenes and films must belong to ctx.  job_ms (may be NULL, njobs entries)
 * receives each job's latency: ms from the start of the batch until its film
 * is final. */
int gr_render_jobs(gr_ctx *ctx, const gr_job *jobs, uint32_t njobs, const gr_batch_params *params,
                   double *job_ms);

/* ---- learned path guide (no reference counterpart) -----------------------
 * A 5-layer MLP (fp16 weights, fp32 accumulation on the matrix cores) maps a
 * surface vertex (position, outgoing direction, shading normal, roughness) to
 * a mixture of 8 von Mises-Fisher lobes over the incident direction.  During a
 * GR_INTEGRATOR_GUIDED trace, a guided vertex whose lobe pick falls in the
 * guided lobe set draws its direction from the mixture with probability alpha
 * and from the picked lobe otherwise, weighted by the one-sample MIS density
 * alpha*q + (1-alpha)*(mean of the set's lobe sampling densities), so the
 * render keeps the plain PT+NEE mean for any network;
 * after each traced batch the network takes Adam steps on that batch's
 * (vertex, direction, f*cos*Li/pdf) records, minimising the KL divergence from
 * the product of BSDF and incident radiance to the mixture. */
typedef struct {
    uint64_t seed;            /* weight initialisation and training-batch draws */
    float learning_rate;      /* Adam step size */
    float alpha;              /* probability of sampling the guide at a guided vertex */
    uint32_t batch;           /* training records per Adam step (multiple of 16) */
    uint32_t steps_per_pass;  /* Adam steps after each traced batch of iterations */
    uint32_t train_paths;     /* paths per traced batch that record training data */
    uint32_t guided_depth;    /* vertices at depth < guided_depth are guided; 0 = all */
    uint32_t warmup_steps;    /* Adam steps taken before the guide is sampled */
    uint32_t lobes;           /* GR_GUIDE_LOBES_*: the lobe picks the guide may redirect */
    uint32_t lobe_select;     /* GR_GUIDE_SELECT_*: how a guided trace picks a vertex's lobe */
    float select_mix;         /* (0, 1]: share of the uniform choice in the lobe-choice probabilities */
    float rr_threshold;       /* >= 0: from the second vertex on, a guided path whose throughput
                                 luminance is below it continues with probability luminance /
                                 rr_threshold, its throughput divided by that (0: no such roulette) */
} gr_guide_params;
/* Guided lobe sets.  CONSISTENT: Disney diffuse, GGX reflection and perfect
 * diffuse, whose reported pdfs are their sampling densities.  ALL: every lobe
 * but the mirror (a delta), clearcoat and transmission included, weighted with
 * the densities their samplers realise (the reported pdfs of those two are
 * not: the GTR1 pdf lacks cos(theta_h), the transmission pdf leaves its half
 * vector unnormalised).  DIFFUSE: the diffuse lobes only. */
#define GR_GUIDE_LOBES_CONSISTENT 0u
#define GR_GUIDE_LOBES_ALL 1u
#define GR_GUIDE_LOBES_DIFFUSE 2u /* Disney diffuse and perfect diffuse: the broad lobes */
/* Lobe choice at every vertex of a guided trace.  UNIFORM: the reference's
 * (each of the n lobes with probability 1/n).  ALBEDO: lobe k with probability
 * c_k = select_mix / n + (1 - select_mix) a_k / sum a, a_k the luminance of
 * its colour times its Fresnel factor at the incoming angle.  LEARNED: the
 * same with a_k the mean contribution lobe k's samples brought on earlier
 * batches of this guide and scene, per primitive and incoming-angle bucket
 * (ALBEDO until every lobe there has 64 samples).  A chosen lobe's sample is
 * divided by n c_k, so each lobe's contribution -- and the plain mean -- is
 * kept. */
#define GR_GUIDE_SELECT_UNIFORM 0u
#define GR_GUIDE_SELECT_ALBEDO 1u
#define GR_GUIDE_SELECT_LEARNED 2u
#define GR_GUIDE_PARAMS 27936  /* fp32 parameters: W1[208][64] b1 W2[64][64] b2 W3 b3 W4 b4 W5[64][32] b5 */
#define GR_GUIDE_MIX_FLOATS 40 /* per query: 8 lobes x (weight, kappa, mean x, y, z) */
#define GR_GUIDE_RECORD_FLOATS 12 /* training record: 8 raw inputs, direction (3), value */
int gr_guide_create(gr_ctx *ctx, const gr_guide_params *params, gr_guide **out);
void gr_guide_destroy(gr_guide *guide);
/* Host copies of the fp32 parameters (n = GR_GUIDE_PARAMS); set also resets Adam. */
int gr_guide_get_params(gr_guide *guide, float *host, uint32_t n);
int gr_guide_set_params(gr_guide *guide, const float *host, uint32_t n);
int gr_guide_steps(gr_guide *guide, uint64_t *adam_steps);
/* Mixture of n vertices from their 8 raw inputs each (device memory, both 16-byte aligned):
 * 40 floats per vertex, 8 lobes of (weight, kappa, mu.x, mu.y, mu.z); slot 2i holds the
 * network's lobe i and slot 2i + 1 its lobe i + 4 (a mixture's lobe order is immaterial). */
int gr_guide_infer(gr_guide *guide, const float *d_raw, uint32_t n, float *d_mix);
/* Mean over n records (device) of the gradient of -(v / mean v) log q(direction). */
int gr_guide_gradient(gr_guide *guide, const float *d_records, uint32_t n, float *d_grad);
/* `steps` Adam steps on batches drawn with replacement from n records (device). */
int gr_guide_train(gr_guide *guide, const float *d_records, uint32_t n, uint32_t steps);
/* The training records of the last traced batch (GR_GUIDE_RECORD_FLOATS each): at
 * most cap of them are copied to d_out (device memory) and their number is written
 * to *n_out.  Waits for the context's stream. */
int gr_guide_last_records(gr_guide *guide, float *d_out, uint32_t cap, uint32_t *n_out);
/* Test hooks: draw directions from n mixtures with 3 uniforms each (out: x, y, z,
 * pdf), and evaluate mixture densities of given directions. */
int gr_guide_mixture_sample(gr_ctx *ctx, const float *d_mix, const float *d_u, uint32_t n, float *d_out);
int gr_guide_mixture_pdf(gr_ctx *ctx, const float *d_mix, const float *d_dirs, uint32_t n, float *d_pdf);

/* Per-function material evaluation on the device (test hook): n items of 24
 * floats in d_in -- wi(3), wo(3), normal, tangent, binormal, geom_normal (12),
 * u_lobe (uint32 bits), u1, u2, camera_to_light, pad(2); 20 floats out per item
 * -- f_world(3), pdf_world, light_world(3), sample_world emission(3),
 * radiance(3), outgoing(3), pdf, kind, count_lobes(DIFFUSE|GLOSSY), and the
 * density of sample_world's draws at wo (the mean over the lobes of the
 * density each lobe's sampler realises; the mirror contributes none). */
int gr_material_query(gr_scene *scene, uint32_t prim, const float *d_in, uint32_t n, float *d_out);

#ifdef __cplusplus
}
#endif
#endif
