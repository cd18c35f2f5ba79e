// gr_internal.h -- library-internal types: context, scene, host BVH and the
// GPU layout of the tree.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gammaray.h"
#include "gr_math.h"

namespace gr {

// ---- host scene -----------------------------------------------------------
enum PrimType : uint32_t { kPrimMesh = 0, kPrimSphere = 1 };

struct Prim {
    PrimType type;
    float material[GR_MAT_FLOATS];
    // mesh
    std::vector<vec3> verts, normals, uvs;  // uvs: (u, v, 0); the last entry is the zero uv
    std::vector<uint32_t> tris;             // 9 per triangle, sentinels resolved
    uint32_t ntris = 0;
    // sphere
    vec3 center{0, 0, 0};
    float radius = 0;
    uint32_t num_components() const { return type == kPrimMesh ? ntris : 1u; }
};

// Reference-order flattened tree (BvhLinearNode, src/geom/bvh.rs:87-92).
struct FlatNode {
    float lo[3], hi[3];
    uint32_t offset;  // leaf: first ordered component; interior: second child
    uint32_t count;   // 0 = interior
    uint32_t axis;
};

struct FlatTree {
    std::vector<FlatNode> nodes;
    std::vector<uint32_t> comp_prim, comp_index;  // ordered components
    uint32_t max_depth = 0;
};

// Build with the reference's SAH algorithm (src/geom/bvh.rs:182-382).
FlatTree build_reference_bvh(const std::vector<Prim>& prims);

// ---- GPU layout -------------------------------------------------------------
// One 64-byte node per interior node of the reference tree, holding the boxes
// of both children so one fetch serves both child tests.  Child refs: interior
// -> node index; leaf -> kLeafBit | first component slot.  Visiting order and
// every box decision stay those of the reference traversal.
constexpr uint32_t kLeafBit = 0x80000000u;

struct alignas(64) GpuNode {
    float4 q0;  // lo0.x lo0.y lo0.z hi0.x     child 0 = first child (DFS i+1)
    float4 q1;  // hi0.y hi0.z lo1.x lo1.y     child 1 = second child (offset)
    float4 q2;  // lo1.z hi1.x hi1.y hi1.z
    uint4 q3;   // ref0 ref1 axis 0
};
static_assert(sizeof(GpuNode) == 64, "node layout");

// One 48-byte record per ordered component (leaf order).  Triangles store
// c, a - c, b - c (the reference's edge1/edge2, computed with the same f32
// subtraction); spheres store centre and radius.  q2.y holds the tag.
constexpr uint32_t kTagSphere = 1u;
constexpr uint32_t kTagLast = 2u;  // last component of its leaf

struct alignas(16) GpuComp {
    float4 q0;  // tri: c.x c.y c.z e1.x       sphere: c.x c.y c.z r
    float4 q1;  // tri: e1.y e1.z e2.x e2.y
    float4 q2;  // tri: e2.z tag 0 0           sphere: 0 tag 0 0
};
static_assert(sizeof(GpuComp) == 48, "component layout");

struct SceneView {
    const GpuNode* nodes;
    const GpuComp* comps;
    const uint2* comp_ids;  // (prim, comp) per slot
    uint32_t root_ref;
    float root_lo[3], root_hi[3];
    float pad;  // 1 + 2 gamma(3)
    uint32_t simple_boxes;  // every child box finite with lo <= hi on each axis
    unsigned long long* overflow;
    uint32_t* next_ray;  // kRayCounters work counters of the persistent traversal kernels
};

struct GpuLayout {
    std::vector<GpuNode> nodes;
    std::vector<GpuComp> comps;
    std::vector<uint2> comp_ids;
    uint32_t root_ref;
    float root_lo[3], root_hi[3];
    bool simple_boxes;  // every child box finite with lo <= hi on each axis
};
GpuLayout make_gpu_layout(const FlatTree& t, const std::vector<Prim>& prims);

// One scene of a multi-scene traversal: the per-scene part of SceneView.
struct SceneRef {
    const GpuNode* nodes;
    const GpuComp* comps;
    const uint2* comp_ids;
    uint32_t root_ref;
    float root_lo[3], root_hi[3];
    uint32_t simple_boxes;
};

// ---- ray sort (gr_sort.hip) -------------------------------------------------
struct SortScratch {
    size_t cap = 0;  // keys
    void* blob = nullptr;
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;    // 256 digit counts per 4096-key block
    uint32_t* totals = nullptr;  // 256
};
// Rounds a requested key width to the widths the sort uses (0 or 3 + 3m).
uint32_t sort_key_bits(uint32_t bits);
hipError_t reserve_sort(SortScratch& ss, size_t n);
// Bytes of sort scratch for n keys, and the layout of that scratch at `base`
// (ss.blob stays null: the memory is not the sort's to free).
size_t sort_scratch_bytes(size_t n);
void carve_sort(SortScratch& ss, void* base, size_t n);
void release_sort(SortScratch& ss);
// Sorts the first n (= *n_ptr if given, at most n_max) rays or segments of
// `in` by (octant, origin Morton code); *perm receives the ray indices in
// sorted order, or nullptr when bits == 0.  ss must hold n_max keys.
hipError_t sort_rays(const SceneView& sv, const float4* in, uint32_t n_max, const uint32_t* n_ptr, bool segments,
                     uint32_t bits, SortScratch& ss, const uint32_t** perm, hipStream_t s);
// Multi-scene form: the key's top bits are the ray's scene, the origin is
// quantised to that scene's box (device table of nscenes entries).
hipError_t sort_rays_multi(const SceneRef* scenes, uint32_t nscenes, const float4* in, uint32_t n_max,
                           const uint32_t* n_ptr, bool segments, uint32_t bits, SortScratch& ss,
                           const uint32_t** perm, hipStream_t s);

// ---- launchers (gr_traverse.hip) -------------------------------------------
// Settings of the traversal launch (gr_tuning).
constexpr uint32_t kDefaultRayChunk = 128;
struct TraceTuning {
    uint32_t sort_bits = 0;
    uint32_t lds_stack = 8;  // 8 waves per SIMD (deeper LDS stacks cap occupancy lower)
    uint32_t waves_per_cu = 0;
    uint32_t refill = 32;      // closest-hit kernels
    uint32_t vis_refill = 40;  // visibility kernels (their segments end at the first occluder)
    uint32_t xcd_ranges = 1;
    uint32_t ray_chunk = kDefaultRayChunk;  // rays a wave reserves per atomic on the work counter
};
// The integrators' traversal setting until one is set or tuned: their bounce
// rays and shadow segments come in queues that shrink bounce by bounce, where
// earlier refills and smaller chunks keep more lanes busy.  Measured on the
// 1080p showcase frame at 64 spp: PT+NEE 667 -> 688 and BDPT 335 -> 344
// Mpaths/s; on one batch of 2^24 incoherent rays and 2^24 segments the same
// setting is 13 % slower than the default above, which ray batches keep.
inline TraceTuning path_tuning_default() {
    TraceTuning t;
    t.refill = 16;
    t.vis_refill = 16;
    t.ray_chunk = 64;
    return t;
}
// Ray counters of the persistent traversal: one per XCD range, each on its own
// 128-byte line (atomics on one line serialise in one L2 channel).
constexpr uint32_t kRayCounters = 8;
constexpr uint32_t kCounterStride = 32;  // uint32 words between counters
constexpr size_t kRayCounterBytes = (size_t)kRayCounters * kCounterStride * sizeof(uint32_t);

// Device-side error counters of a context (gr_stats, in this order).
enum StatIndex : int { kStatNonfiniteBsdf = 0, kStatNonfiniteRadiance = 1, kStatBadSplats = 2, kStats = 4 };

// Everything a traversal launch needs besides its rays.  With `scenes` set,
// every ray names its scene in the pad word after its origin (an index into
// `scenes`, nscenes entries); sv then only supplies the work counter, the
// overflow counter and the slab padding.
struct TraceSetup {
    SceneView sv;
    TraceTuning tu;
    SortScratch* sort;       // holds >= n_max keys when tu.sort_bits != 0
    gr_ctx* timing = nullptr;  // set when kernel timing is on
    const SceneRef* scenes = nullptr;
    uint32_t nscenes = 0;
};

// Kernel timing (gr_api.cpp): brackets launches with events on stream s.
hipEvent_t timing_begin(gr_ctx* ctx, hipStream_t s);
void timing_end(gr_ctx* ctx, int kind, hipEvent_t begin, hipStream_t s);

// n_ptr: if not null, the ray count is read on the device from *n_ptr (<= n).
hipError_t launch_intersect(const TraceSetup& ts, const gr_ray* rays, uint32_t n, const uint32_t* n_ptr,
                            gr_hit* hits, uint32_t* counts, hipStream_t s);
hipError_t launch_occluded(const TraceSetup& ts, const gr_segment* segs, uint32_t n, const uint32_t* n_ptr,
                           uint32_t* visible, int mode, uint32_t* counts, hipStream_t s);
// Resident waves per CU the traversal kernel reaches with this LDS stack depth.
int trace_occupancy(uint32_t lds_stack);

// ---- shading data ------------------------------------------------------------
// Per prim: where its mesh data lives, or its sphere, plus its area.
struct GpuPrim {
    uint32_t type, tri_off, vert_off, norm_off;
    uint32_t uv_off, ntris, cdf_off, pad0;
    float center[3], radius;
    float area, pad1, pad2, pad3;
};
static_assert(sizeof(GpuPrim) == 64, "prim layout");

struct GpuMaterial;

struct ShadeView {
    const GpuMaterial* mats;
    const GpuPrim* prims;
    const float4* verts;
    const float4* normals;
    const float2* uvs;
    const uint32_t* tris;  // 9 per triangle, sentinels resolved
    const float* cdf;      // per mesh: area CDF over its triangles
    const uint32_t* lights;
    uint32_t num_lights;
};

}  // namespace gr

// ---- ABI objects ----------------------------------------------------------
struct gr_ctx {
    // One reference for the caller's handle and one per live scene, film and
    // guide: gr_ctx_destroy drops the handle's, and the context is freed when
    // the last reference goes, so objects that outlive the handle stay valid.
    std::atomic<int> refs{1};
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    bool timing = false;
    unsigned long long* d_stats = nullptr;  // gr::kStats error counters
    struct TimedLaunch {
        int kind;
        hipEvent_t begin, end;
    };
    std::vector<TimedLaunch> timed;
};

struct gr_scene {
    gr_ctx* ctx = nullptr;
    std::vector<gr::Prim> prims;
    bool committed = false;
    gr::FlatTree tree;
    uint32_t num_interior = 0, num_lights = 0;
    double build_seconds = 0;
    gr::GpuNode* d_nodes = nullptr;
    gr::GpuComp* d_comps = nullptr;
    uint2* d_comp_ids = nullptr;
    unsigned long long* d_overflow = nullptr;
    uint32_t* d_next_ray = nullptr;
    gr::SceneView view{};
    gr::TraceTuning tuning{};                                       // ray batches (gr_intersect_*, gr_visibility)
    gr::TraceTuning render_tuning = gr::path_tuning_default();  // the integrators (gr_trace, gr_render_jobs)
    gr::SortScratch sort{};
    // shading data (gr_render.hip)
    void* d_shade_blob = nullptr;  // one allocation holding every shading array
    gr::ShadeView shade{};
};

struct gr_guide {
    gr_ctx* ctx = nullptr;
    gr_guide_params prm{};
    void* blob = nullptr;  // parameters, Adam moments, gradient, fp16 fragments, counters
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *fwd = nullptr, *bwd = nullptr;  // fp16 MFMA fragments of the weights (K=16)
    uint16_t* fwd32 = nullptr;                // inference fragments (K=32)
    float* stats = nullptr;                   // [0] sum of record values
    uint32_t* count = nullptr;                // records available for training
    uint64_t steps = 0;                       // Adam steps taken
    // dW operands of one training batch, grown on demand, with the gradient's
    // partial sums: per-tile bias sums and one full gradient per split of the
    // batch, summed in a fixed order (no float atomics: training is reproducible)
    void* work = nullptr;
    uint32_t work_tiles = 0;
    uint16_t *at = nullptr, *dz = nullptr;
    float *bias_part = nullptr, *part = nullptr;
    // training records of one traced batch (compacted in slot order), grown on
    // demand, with the compaction's per-block counts and offsets
    void* recs_blob = nullptr;
    size_t recs_cap = 0;
    float* recs = nullptr;
    uint32_t *rec_bcount = nullptr, *rec_boff = nullptr;
    // per-block partial sums of record values, grown on demand
    float* vpart = nullptr;
    size_t vpart_cap = 0;
    // learned lobe choice (GR_GUIDE_SELECT_LEARNED): per (primitive, incoming
    // bucket, lobe) fixed-point sums of the chosen lobe's contribution and
    // their counts, and the choice probabilities made from them, for the
    // scene they were gathered on
    void* lobe_blob = nullptr;
    unsigned long long* lobe_sum = nullptr;
    uint32_t* lobe_cnt = nullptr;
    float4* lobe_c = nullptr;
    float* lobe_m = nullptr;
    size_t lobe_prims = 0;
    const gr_scene* lobe_scene = nullptr;
};

namespace gr {
// gr_guide.hip
// Mixture (kMixFloats each) of the first n (= *n_ptr if given, <= n_max) raw inputs.
hipError_t guide_infer(gr_guide* g, const float* raw, uint32_t n_max, const uint32_t* n_ptr, float* mix,
                       hipStream_t s);
// Sum of the values of the first n records into g->stats[0] (which the caller zeroes).
hipError_t guide_record_stats(gr_guide* g, const float* recs, uint32_t n_max, const uint32_t* n_ptr, hipStream_t s);
hipError_t guide_zero_stats(gr_guide* g, hipStream_t s);
// `steps` Adam steps on batches of g->prm.batch records drawn from the first n records.
hipError_t guide_adam_steps(gr_guide* g, const float* recs, uint32_t n_max, const uint32_t* n_ptr, uint32_t steps,
                            hipStream_t s);
hipError_t guide_reserve_records(gr_guide* g, size_t n);
}  // namespace gr

// Per-film wavefront buffers, grown on demand by gr_trace.
struct RenderBuffers {
    size_t capacity = 0;  // paths
    void* blob = nullptr;
    float4 *rays[2] = {nullptr, nullptr};   // 2 float4 per ray (gr_ray)
    uint32_t *queue[2] = {nullptr, nullptr};
    uint4* hits = nullptr;
    float4 *T = nullptr, *L = nullptr, *aux = nullptr;
    float2* st = nullptr;
    float4* seg = nullptr;      // shadow segments, 2 float4 each
    float4* contrib = nullptr;  // shadow contributions
    uint32_t* seg_path = nullptr;
    uint32_t* vis = nullptr;
    uint32_t* counts = nullptr;  // [0..1] queue lengths, [2] shadow queue length
    // BDPT vertex storage
    size_t bdpt_capacity = 0;
    void* bdpt_blob = nullptr;
    // guided tracing: queries, mixtures, per-path query slots, training vertex records
    size_t guide_capacity = 0, guide_vrecs = 0;
    void* guide_blob = nullptr;
    float *graw = nullptr, *gmix = nullptr;
    uint32_t* gslot = nullptr;
    float4* vrec = nullptr;
};

struct gr_film {
    gr_ctx* ctx = nullptr;
    uint32_t width = 0, height = 0;
    float4* d_pixels = nullptr;  // (accum r, g, b, weight)
    RenderBuffers rb;
    // asynchronous readback: a device snapshot, pinned host copy, a side stream
    float4* d_snapshot = nullptr;
    float4* h_pinned = nullptr;
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_ready = nullptr, copy_done = nullptr;
    bool copy_pending = false;
};

namespace gr {
// Context references held by scenes, films and guides (gr_api.cpp).
void ctx_retain(gr_ctx* ctx);
void ctx_release(gr_ctx* ctx);
inline TraceSetup trace_setup(gr_scene* s) {
    return TraceSetup{s->view, s->tuning, &s->sort, s->ctx->timing ? s->ctx : nullptr};
}
// the integrators' launches: the scene's render setting
inline TraceSetup render_setup(gr_scene* s) {
    return TraceSetup{s->view, s->render_tuning, &s->sort, s->ctx->timing ? s->ctx : nullptr};
}
// gr_render.hip
hipError_t upload_shading(gr_scene* s, std::vector<GpuPrim>* prims_out);
hipError_t render_trace(gr_scene* sc, const gr_camera& cam, const gr_trace_params& prm, gr_film* film);
hipError_t render_jobs(gr_ctx* ctx, const gr_job* jobs, uint32_t njobs, const gr_batch_params& prm, double* job_ms);
hipError_t material_query(gr_scene* sc, uint32_t prim, const float* in, uint32_t n, float* out);
void release_render_buffers(RenderBuffers& rb);
// gr_noalloc.hip: resolves the no-allocation scratch array's address on the
// current device (once per device; called by gr_ctx_create).
hipError_t noalloc_setup(int device);
// gr_exr.cpp
bool write_exr(const float* rgbw, uint32_t width, uint32_t height, const char* path);
void encode_exr(const float* rgbw, uint32_t width, uint32_t height, std::vector<uint8_t>* out);
}  // namespace gr
