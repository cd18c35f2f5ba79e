// gr_api.cpp -- the extern "C" boundary (include/gammaray.h).
//
// Ownership mirrors the reference: a scene takes copies of its prims and is
// immutable after commit (Stage owns Bvh owns prims, src/render/stage.rs:10-21);
// the film is a separate object borrowed by each trace call.  No call aborts
// or throws across the boundary: failures return a negative code and leave a
// message in gr_ctx_last_error().
#pragma clang fp contract(off)

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <new>

#include "gr_internal.h"

using namespace gr;

namespace {

int fail(gr_ctx* ctx, int code, const char* fmt, const char* detail = nullptr) {
    if (ctx) {
        char buf[512];
        std::snprintf(buf, sizeof(buf), fmt, detail ? detail : "");
        ctx->err = buf;
    }
    return code;
}

int hip_fail(gr_ctx* ctx, hipError_t e, const char* where) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
    if (ctx) ctx->err = buf;
    return GR_ERR_HIP;
}

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& src, hipStream_t s) {
    *dst = nullptr;
    size_t bytes = src.size() * sizeof(T);
    if (bytes == 0) bytes = sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), bytes);
    if (e != hipSuccess) return e;
    if (!src.empty()) e = hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s);
    return e;
}

void free_device(gr_scene* s) {
    (void)hipFree(s->d_nodes);
    (void)hipFree(s->d_comps);
    (void)hipFree(s->d_comp_ids);
    (void)hipFree(s->d_overflow);
    (void)hipFree(s->d_next_ray);
    (void)hipFree(s->d_shade_blob);
    gr::release_sort(s->sort);
    s->d_shade_blob = nullptr;
    s->shade = gr::ShadeView{};
    s->d_next_ray = nullptr;
    s->d_nodes = nullptr;
    s->d_comps = nullptr;
    s->d_comp_ids = nullptr;
    s->d_overflow = nullptr;
}

}  // namespace

namespace gr {
hipEvent_t timing_begin(gr_ctx*, hipStream_t s) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    if (hipEventRecord(e, s) != hipSuccess) {
        (void)hipEventDestroy(e);
        return nullptr;
    }
    return e;
}

void timing_end(gr_ctx* ctx, int kind, hipEvent_t begin, hipStream_t s) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess || hipEventRecord(e, s) != hipSuccess) {
        if (e) (void)hipEventDestroy(e);
        (void)hipEventDestroy(begin);
        return;
    }
    ctx->timed.push_back(gr_ctx::TimedLaunch{kind, begin, e});
}
}  // namespace gr

static void drop_timed(gr_ctx* ctx) {
    for (auto& t : ctx->timed) {
        (void)hipEventDestroy(t.begin);
        (void)hipEventDestroy(t.end);
    }
    ctx->timed.clear();
}

namespace gr {
void ctx_retain(gr_ctx* ctx) { ctx->refs.fetch_add(1, std::memory_order_relaxed); }

void ctx_release(gr_ctx* ctx) {
    if (ctx->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    if (ctx->device != GR_HOST_ONLY) {
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        drop_timed(ctx);
        if (ctx->own_stream) {
            (void)hipStreamSynchronize(ctx->own_stream);
            (void)hipStreamDestroy(ctx->own_stream);
        }
        (void)hipFree(ctx->d_stats);
    }
    delete ctx;
}
}  // namespace gr

extern "C" {

int gr_ctx_create(int device, gr_ctx** out) {
    if (!out) return GR_ERR_INVALID;
    *out = nullptr;
    if (device == GR_HOST_ONLY) {
        gr_ctx* c = new (std::nothrow) gr_ctx();
        if (!c) return GR_ERR_NOMEM;
        c->device = GR_HOST_ONLY;
        *out = c;
        return GR_OK;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return GR_ERR_INVALID;
    gr_ctx* c = new (std::nothrow) gr_ctx();
    if (!c) return GR_ERR_NOMEM;
    c->device = device;
    hipError_t e = hipSetDevice(device);
    // A blocking stream: ordered with work on the legacy default stream, where
    // a caller's uploads often run.
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamDefault);
    if (e == hipSuccess) e = gr::noalloc_setup(device);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->d_stats), gr::kStats * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemsetAsync(c->d_stats, 0, gr::kStats * sizeof(unsigned long long), c->own_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->own_stream);
    if (e != hipSuccess) {
        (void)hipFree(c->d_stats);
        if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
        delete c;
        return GR_ERR_HIP;
    }
    c->stream = c->own_stream;
    *out = c;
    return GR_OK;
}

void gr_ctx_destroy(gr_ctx* ctx) {
    if (!ctx) return;
    gr::ctx_release(ctx);
}

int gr_ctx_set_stream(gr_ctx* ctx, void* stream) {
    if (!ctx) return GR_ERR_INVALID;
    ctx->stream = reinterpret_cast<hipStream_t>(stream);
    return GR_OK;
}

int gr_ctx_use_own_stream(gr_ctx* ctx) {
    if (!ctx) return GR_ERR_INVALID;
    ctx->stream = ctx->own_stream;
    return GR_OK;
}

void* gr_ctx_stream(gr_ctx* ctx) { return ctx ? reinterpret_cast<void*>(ctx->stream) : nullptr; }

int gr_ctx_synchronize(gr_ctx* ctx) {
    if (!ctx) return GR_ERR_INVALID;
    if (ctx->device == GR_HOST_ONLY) return GR_OK;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? GR_OK : hip_fail(ctx, e, "hipStreamSynchronize");
}

const char* gr_ctx_last_error(gr_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int gr_ctx_set_kernel_timing(gr_ctx* ctx, int enable) {
    if (!ctx) return GR_ERR_INVALID;
    if (ctx->device == GR_HOST_ONLY) return fail(ctx, GR_ERR_STATE, "host-only context%s");
    ctx->timing = enable != 0;
    return GR_OK;
}

int gr_ctx_take_kernel_times(gr_ctx* ctx, double* ms, uint64_t* launches) {
    if (!ctx) return GR_ERR_INVALID;
    if (ctx->device == GR_HOST_ONLY) return fail(ctx, GR_ERR_STATE, "host-only context%s");
    double acc[GR_KERNEL_KINDS] = {0.0};
    uint64_t cnt[GR_KERNEL_KINDS] = {0};
    hipError_t e = hipSuccess;
    for (auto& t : ctx->timed) {
        if (e == hipSuccess) e = hipEventSynchronize(t.end);
        float v = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&v, t.begin, t.end);
        if (e == hipSuccess && t.kind >= 0 && t.kind < GR_KERNEL_KINDS) {
            acc[t.kind] += v;
            cnt[t.kind] += 1;
        }
    }
    drop_timed(ctx);
    if (e != hipSuccess) return hip_fail(ctx, e, "kernel timing");
    for (int k = 0; k < GR_KERNEL_KINDS; ++k) {
        if (ms) ms[k] = acc[k];
        if (launches) launches[k] = cnt[k];
    }
    return GR_OK;
}

int gr_ctx_take_stats(gr_ctx* ctx, gr_stats* out) {
    if (!ctx || !out) return GR_ERR_INVALID;
    if (ctx->device == GR_HOST_ONLY) return fail(ctx, GR_ERR_STATE, "host-only context%s");
    unsigned long long v[gr::kStats] = {0};
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess)
        e = hipMemcpyAsync(v, ctx->d_stats, sizeof(v), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->d_stats, 0, sizeof(v), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "stats readback");
    std::memset(out, 0, sizeof(*out));
    out->nonfinite_bsdf_samples = v[gr::kStatNonfiniteBsdf];
    out->nonfinite_radiance = v[gr::kStatNonfiniteRadiance];
    out->bad_splats = v[gr::kStatBadSplats];
    return GR_OK;
}

int gr_scene_create(gr_ctx* ctx, gr_scene** out) {
    if (!ctx || !out) return GR_ERR_INVALID;
    gr_scene* s = new (std::nothrow) gr_scene();
    if (!s) return fail(ctx, GR_ERR_NOMEM, "out of host memory%s");
    s->ctx = ctx;
    gr::ctx_retain(ctx);
    *out = s;
    return GR_OK;
}

void gr_scene_destroy(gr_scene* s) {
    if (!s) return;
    gr_ctx* ctx = s->ctx;
    if (ctx->device != GR_HOST_ONLY) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        free_device(s);
    }
    delete s;
    gr::ctx_release(ctx);
}

int gr_scene_add_mesh(gr_scene* s, const float* verts, uint32_t nverts, const float* normals,
                      uint32_t nnormals, const float* uvs, uint32_t nuvs, const uint32_t* tris,
                      uint32_t ntris, const float* material, uint32_t* prim_out) {
    if (!s) return GR_ERR_INVALID;
    gr_ctx* ctx = s->ctx;
    if (s->committed) return fail(ctx, GR_ERR_STATE, "scene already committed%s");
    if (!verts || !tris || !material || nverts == 0 || ntris == 0)
        return fail(ctx, GR_ERR_INVALID, "mesh needs vertices, triangles and a material%s");
    if ((nnormals && !normals) || (nuvs && !uvs)) return fail(ctx, GR_ERR_INVALID, "null normal/uv array%s");
    for (uint64_t i = 0; i < (uint64_t)ntris * 9; ++i) {
        const uint32_t k = (uint32_t)(i % 9), v = tris[i];
        if ((k < 3 && v >= nverts) || (k >= 3 && k < 6 && v != kNone && v >= nnormals) ||
            (k >= 6 && v != kNone && v >= nuvs))
            return fail(ctx, GR_ERR_INVALID, "triangle index out of range%s");
    }
    Prim p;
    p.type = kPrimMesh;
    std::memcpy(p.material, material, sizeof(p.material));
    p.verts.resize(nverts);
    for (uint32_t i = 0; i < nverts; ++i) p.verts[i] = mk(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    p.normals.resize(nnormals);
    for (uint32_t i = 0; i < nnormals; ++i)
        p.normals[i] = mk(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
    p.uvs.resize(nuvs + 1);
    for (uint32_t i = 0; i < nuvs; ++i) p.uvs[i] = mk(uvs[2 * i], uvs[2 * i + 1], 0.0f);
    p.uvs[nuvs] = mk(0.0f, 0.0f, 0.0f);  // the zero uv of src/geom/mesh.rs:133-136
    p.ntris = ntris;
    p.tris.assign(tris, tris + 9 * (size_t)ntris);
    for (uint32_t t = 0; t < ntris; ++t) {
        uint32_t* tr = &p.tris[9 * (size_t)t];
        // Missing or zero-length shading normals -> geometric normal
        // (src/geom/mesh.rs:138-153); missing uvs -> the zero uv (:129-137).
        bool has_n = tr[3] != kNone && tr[4] != kNone && tr[5] != kNone;
        if (has_n) {
            for (int k = 3; k < 6; ++k)
                if (nearly_zero(dot(p.normals[tr[k]], p.normals[tr[k]]))) has_n = false;
        }
        if (!has_n) tr[3] = tr[4] = tr[5] = kNone;
        if (tr[6] == kNone || tr[7] == kNone || tr[8] == kNone) tr[6] = tr[7] = tr[8] = nuvs;
    }
    s->prims.push_back(std::move(p));
    if (prim_out) *prim_out = (uint32_t)(s->prims.size() - 1);
    return GR_OK;
}

int gr_scene_add_sphere(gr_scene* s, const float center[3], float radius, const float* material,
                        uint32_t* prim_out) {
    if (!s) return GR_ERR_INVALID;
    if (s->committed) return fail(s->ctx, GR_ERR_STATE, "scene already committed%s");
    if (!center || !material) return fail(s->ctx, GR_ERR_INVALID, "sphere needs a centre and a material%s");
    Prim p;
    p.type = kPrimSphere;
    std::memcpy(p.material, material, sizeof(p.material));
    p.center = mk(center[0], center[1], center[2]);
    p.radius = radius;
    s->prims.push_back(std::move(p));
    if (prim_out) *prim_out = (uint32_t)(s->prims.size() - 1);
    return GR_OK;
}

int gr_scene_commit(gr_scene* s) {
    if (!s) return GR_ERR_INVALID;
    gr_ctx* ctx = s->ctx;
    if (s->committed) return fail(ctx, GR_ERR_STATE, "scene already committed%s");
    if (s->prims.empty()) return fail(ctx, GR_ERR_INVALID, "scene has no prims%s");
    size_t total = 0;
    for (const Prim& p : s->prims) total += p.num_components();
    if (total >= (size_t)kLeafBit) return fail(ctx, GR_ERR_INVALID, "too many components%s");
    auto t0 = std::chrono::steady_clock::now();
    try {
        s->tree = build_reference_bvh(s->prims);
    } catch (const std::bad_alloc&) {
        return fail(ctx, GR_ERR_NOMEM, "out of host memory building the BVH%s");
    }
    GpuLayout lay = make_gpu_layout(s->tree, s->prims);
    s->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    s->num_interior = (uint32_t)lay.nodes.size();
    s->num_lights = 0;
    for (const Prim& p : s->prims)
        if ((int)p.material[0] == 1) s->num_lights++;
    if (ctx->device == GR_HOST_ONLY) {
        (void)upload_shading(s, nullptr);  // builds the material table; uploads nothing
        s->committed = true;
        return GR_OK;
    }

    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = upload_shading(s, nullptr);
    if (e == hipSuccess) e = upload(&s->d_nodes, lay.nodes, ctx->stream);
    if (e == hipSuccess) e = upload(&s->d_comps, lay.comps, ctx->stream);
    if (e == hipSuccess) e = upload(&s->d_comp_ids, lay.comp_ids, ctx->stream);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_overflow), sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemsetAsync(s->d_overflow, 0, sizeof(unsigned long long), ctx->stream);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->d_next_ray), gr::kRayCounterBytes);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        free_device(s);
        return hip_fail(ctx, e, "scene upload");
    }
    SceneView& v = s->view;
    v.nodes = s->d_nodes;
    v.comps = s->d_comps;
    v.comp_ids = s->d_comp_ids;
    v.root_ref = lay.root_ref;
    for (int k = 0; k < 3; ++k) {
        v.root_lo[k] = lay.root_lo[k];
        v.root_hi[k] = lay.root_hi[k];
    }
    v.pad = slab_pad();
    v.simple_boxes = lay.simple_boxes ? 1u : 0u;
    v.overflow = s->d_overflow;
    v.next_ray = s->d_next_ray;
    s->committed = true;
    return GR_OK;
}

int gr_scene_info_get(gr_scene* s, gr_scene_info* out) {
    if (!s || !out) return GR_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    out->num_prims = (uint32_t)s->prims.size();
    out->num_components = (uint32_t)s->tree.comp_prim.size();
    out->num_nodes = (uint32_t)s->tree.nodes.size();
    out->num_interior = s->num_interior;
    out->num_lights = s->num_lights;
    out->max_depth = s->tree.max_depth;
    out->node_bytes = (uint64_t)s->num_interior * sizeof(GpuNode);
    out->component_bytes = (uint64_t)out->num_components * (sizeof(GpuComp) + sizeof(uint2));
    out->build_seconds = s->build_seconds;
    return s->committed ? GR_OK : fail(s->ctx, GR_ERR_STATE, "scene not committed%s");
}

int gr_scene_export_bvh(gr_scene* s, float* bbox6, uint32_t* offset, uint32_t* count,
                        uint32_t* axis, uint32_t* comp_prim, uint32_t* comp_index) {
    if (!s) return GR_ERR_INVALID;
    if (!s->committed) return fail(s->ctx, GR_ERR_STATE, "scene not committed%s");
    const auto& nodes = s->tree.nodes;
    for (size_t i = 0; i < nodes.size(); ++i) {
        if (bbox6) {
            for (int k = 0; k < 3; ++k) {
                bbox6[6 * i + k] = nodes[i].lo[k];
                bbox6[6 * i + 3 + k] = nodes[i].hi[k];
            }
        }
        if (offset) offset[i] = nodes[i].offset;
        if (count) count[i] = nodes[i].count;
        if (axis) axis[i] = nodes[i].axis;
    }
    for (size_t i = 0; i < s->tree.comp_prim.size(); ++i) {
        if (comp_prim) comp_prim[i] = s->tree.comp_prim[i];
        if (comp_index) comp_index[i] = s->tree.comp_index[i];
    }
    return GR_OK;
}

int gr_intersect_closest(gr_scene* s, const gr_ray* d_rays, uint32_t n, gr_hit* d_hits,
                         uint32_t* d_counts) {
    if (!s) return GR_ERR_INVALID;
    if (!s->committed) return fail(s->ctx, GR_ERR_STATE, "scene not committed%s");
    if (s->ctx->device == GR_HOST_ONLY) return fail(s->ctx, GR_ERR_STATE, "host-only context%s");
    if (n && (!d_rays || !d_hits)) return fail(s->ctx, GR_ERR_INVALID, "null ray/hit buffer%s");
    hipError_t e = s->tuning.sort_bits ? gr::reserve_sort(s->sort, n) : hipSuccess;
    if (e == hipSuccess) e = launch_intersect(gr::trace_setup(s), d_rays, n, nullptr, d_hits, d_counts, s->ctx->stream);
    return e == hipSuccess ? GR_OK : hip_fail(s->ctx, e, "intersect launch");
}

int gr_occluded(gr_scene* s, const gr_segment* d_segs, uint32_t n, uint32_t* d_visible, int mode,
                uint32_t* d_counts) {
    if (!s) return GR_ERR_INVALID;
    if (!s->committed) return fail(s->ctx, GR_ERR_STATE, "scene not committed%s");
    if (mode != GR_OCCLUSION_EXACT && mode != GR_OCCLUSION_ANYHIT)
        return fail(s->ctx, GR_ERR_INVALID, "unknown occlusion mode%s");
    if (s->ctx->device == GR_HOST_ONLY) return fail(s->ctx, GR_ERR_STATE, "host-only context%s");
    if (n && (!d_segs || !d_visible)) return fail(s->ctx, GR_ERR_INVALID, "null segment/output buffer%s");
    hipError_t e = s->tuning.sort_bits ? gr::reserve_sort(s->sort, n) : hipSuccess;
    if (e == hipSuccess)
        e = launch_occluded(gr::trace_setup(s), d_segs, n, nullptr, d_visible, mode, d_counts, s->ctx->stream);
    return e == hipSuccess ? GR_OK : hip_fail(s->ctx, e, "occlusion launch");
}

static bool valid_tuning(const gr_tuning& t) {
    return t.sort_bits <= 30 && (t.lds_stack == 8 || t.lds_stack == 16 || t.lds_stack == 32) && t.refill >= 1 &&
           t.refill <= 64 && t.waves_per_cu <= 64 && t.xcd_ranges <= 1 && t.ray_chunk <= 65536 &&
           t.vis_refill <= 64;
}

int gr_scene_set_tuning(gr_scene* s, const gr_tuning* t) {
    if (!s || !t) return GR_ERR_INVALID;
    if (!valid_tuning(*t)) return fail(s->ctx, GR_ERR_INVALID, "tuning out of range%s");
    s->tuning.sort_bits = gr::sort_key_bits(t->sort_bits);
    s->tuning.lds_stack = t->lds_stack;
    s->tuning.waves_per_cu = t->waves_per_cu;
    s->tuning.refill = t->refill;
    s->tuning.xcd_ranges = t->xcd_ranges;
    s->tuning.ray_chunk = t->ray_chunk ? t->ray_chunk : gr::kDefaultRayChunk;
    s->tuning.vis_refill = t->vis_refill ? t->vis_refill : t->refill;
    s->render_tuning = s->tuning;  // a setting given applies to every launch on the scene
    return GR_OK;
}

int gr_scene_get_tuning(gr_scene* s, gr_tuning* out) {
    if (!s || !out) return GR_ERR_INVALID;
    out->sort_bits = s->tuning.sort_bits;
    out->lds_stack = s->tuning.lds_stack;
    out->waves_per_cu = s->tuning.waves_per_cu;
    out->refill = s->tuning.refill;
    out->xcd_ranges = s->tuning.xcd_ranges;
    out->ray_chunk = s->tuning.ray_chunk;
    out->vis_refill = s->tuning.vis_refill;
    return GR_OK;
}

namespace {
// The fastest of 6 timed passes (after one untimed) of the workload under
// setting tu, launched with `base` (a scene's view, or every job's scenes).
hipError_t time_workload(const gr::TraceSetup& base, const gr::TraceTuning& tu, hipStream_t st, const gr_ray* rays,
                         uint32_t n, gr_hit* hits, const gr_segment* segs, uint32_t m, uint32_t* vis, int mode,
                         hipEvent_t a, hipEvent_t b, float* best) {
    gr::TraceSetup ts = base;
    ts.tu = tu;
    *best = 3.0e38f;
    for (int rep = 0; rep < 7; ++rep) {
        hipError_t e = hipEventRecord(a, st);
        if (e == hipSuccess && n) e = gr::launch_intersect(ts, rays, n, nullptr, hits, nullptr, st);
        if (e == hipSuccess && m) e = gr::launch_occluded(ts, segs, m, nullptr, vis, mode, nullptr, st);
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e != hipSuccess) return e;
        if (rep > 0 && ms < *best) *best = ms;
    }
    return hipSuccess;
}

// Times candidate settings one knob at a time from the defaults and returns
// the fastest in *out.  A candidate must win by more than 3 % to replace the
// incumbent (run-to-run differences of the same setting reach 1-2 %, and a
// smaller margin let them decide between equal settings).
hipError_t search_tuning(const gr::TraceSetup& base, hipStream_t st, const gr_ray* d_rays, uint32_t n,
                         gr_hit* d_hits, const gr_segment* d_segs, uint32_t m, uint32_t* d_visible, int mode,
                         gr::TraceTuning* out) {
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    gr::TraceTuning best{};  // the defaults: sort off, XCD ranges, 8 LDS entries, occupancy limit,
                            // refill at 32 (40 for visibility), 128-ray chunks
    float best_ms = 0.0f;
    auto timed = [&](const gr::TraceTuning& tu, float* ms) {
        return time_workload(base, tu, st, d_rays, n, d_hits, d_segs, m, d_visible, mode, a, b, ms);
    };
    if (e == hipSuccess) e = timed(best, &best_ms);
    auto consider = [&](gr::TraceTuning cand) {
        if (e != hipSuccess) return;
        float ms = 0.0f;
        e = timed(cand, &ms);
        if (e == hipSuccess && ms < 0.97f * best_ms) {
            best = cand;
            best_ms = ms;
        }
    };
    for (uint32_t bits : {12u, 15u, 18u, 21u, 24u, 30u}) {  // 15 and 18: two and three radix passes
        gr::TraceTuning c = best;
        c.sort_bits = bits;
        consider(c);
    }
    {
        gr::TraceTuning c = best;
        c.xcd_ranges = best.xcd_ranges ? 0u : 1u;
        consider(c);
    }
    for (uint32_t lds : {8u, 16u, 32u}) {
        gr::TraceTuning c = best;
        c.lds_stack = lds;
        if (lds != best.lds_stack) consider(c);
    }
    const int occ = gr::trace_occupancy(best.lds_stack);
    for (int frac : {3, 2, 1}) {
        gr::TraceTuning c = best;
        c.waves_per_cu = (uint32_t)std::max(1, occ * frac / 4);
        if ((int)c.waves_per_cu < occ) consider(c);
    }
    for (uint32_t rf : {24u, 40u, 48u, 56u}) {
        gr::TraceTuning c = best;
        c.refill = rf;
        if (rf != best.refill) consider(c);
    }
    if (m) {
        for (uint32_t rf : {24u, 32u, 48u, 56u}) {
            gr::TraceTuning c = best;
            c.vis_refill = rf;
            if (rf != best.vis_refill) consider(c);
        }
    }
    for (uint32_t ch : {64u, 128u, 512u, 1024u}) {
        gr::TraceTuning c = best;
        c.ray_chunk = ch;
        if (ch != best.ray_chunk) consider(c);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (e == hipSuccess) *out = best;
    return e;
}

void tuning_out(const gr::TraceTuning& t, gr_tuning* out) {
    out->sort_bits = t.sort_bits;
    out->lds_stack = t.lds_stack;
    out->waves_per_cu = t.waves_per_cu;
    out->refill = t.refill;
    out->xcd_ranges = t.xcd_ranges;
    out->ray_chunk = t.ray_chunk;
    out->vis_refill = t.vis_refill;
}
}  // namespace

int gr_scene_autotune_mixed(gr_scene* s, const gr_ray* d_rays, uint32_t n, gr_hit* d_hits, const gr_segment* d_segs,
                            uint32_t m, uint32_t* d_visible, int mode, gr_tuning* out) {
    if (!s) return GR_ERR_INVALID;
    if (!s->committed) return fail(s->ctx, GR_ERR_STATE, "scene not committed%s");
    if (s->ctx->device == GR_HOST_ONLY) return fail(s->ctx, GR_ERR_STATE, "host-only context%s");
    if (n + (uint64_t)m == 0 || (n && (!d_rays || !d_hits)) || (m && (!d_segs || !d_visible)))
        return fail(s->ctx, GR_ERR_INVALID, "autotune needs rays with a hit buffer and/or segments with a result buffer%s");
    if (mode != GR_OCCLUSION_EXACT && mode != GR_OCCLUSION_ANYHIT)
        return fail(s->ctx, GR_ERR_INVALID, "unknown occlusion mode%s");
    hipError_t e = gr::reserve_sort(s->sort, std::max(n, m));
    gr::TraceTuning best{};
    if (e == hipSuccess)
        e = search_tuning(gr::TraceSetup{s->view, best, &s->sort}, s->ctx->stream, d_rays, n, d_hits, d_segs, m,
                          d_visible, mode, &best);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "autotune");
    s->tuning = best;
    s->render_tuning = best;
    return out ? gr_scene_get_tuning(s, out) : GR_OK;
}

int gr_jobs_autotune(gr_ctx* ctx, gr_scene* const* scenes, uint32_t nscenes, const gr_ray* d_rays, uint32_t n,
                     gr_hit* d_hits, const gr_segment* d_segs, uint32_t m, uint32_t* d_visible, gr_tuning* out) {
    if (!ctx || !out) return GR_ERR_INVALID;
    if (ctx->device == GR_HOST_ONLY) return fail(ctx, GR_ERR_STATE, "host-only context%s");
    if (!scenes || nscenes == 0) return fail(ctx, GR_ERR_INVALID, "no scenes%s");
    for (uint32_t k = 0; k < nscenes; ++k) {
        if (!scenes[k] || !scenes[k]->committed) return fail(ctx, GR_ERR_STATE, "scene not committed%s");
        if (scenes[k]->ctx != ctx) return fail(ctx, GR_ERR_INVALID, "scene of another context%s");
    }
    if (n + (uint64_t)m == 0 || (n && (!d_rays || !d_hits)) || (m && (!d_segs || !d_visible)))
        return fail(ctx, GR_ERR_INVALID, "autotune needs rays with a hit buffer and/or segments with a result buffer%s");
    // the scenes as the batch's traversal sees them
    std::vector<gr::SceneRef> hs(nscenes);
    for (uint32_t k = 0; k < nscenes; ++k) {
        const gr::SceneView& v = scenes[k]->view;
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
    void* d_refs = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = gr::reserve_sort(scenes[0]->sort, std::max(n, m));
    if (e == hipSuccess) e = hipMalloc(&d_refs, hs.size() * sizeof(gr::SceneRef));
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_refs, hs.data(), hs.size() * sizeof(gr::SceneRef), hipMemcpyHostToDevice, ctx->stream);
    gr::TraceTuning best{};
    if (e == hipSuccess)
        e = search_tuning(gr::TraceSetup{scenes[0]->view, best, &scenes[0]->sort, nullptr,
                                         static_cast<const gr::SceneRef*>(d_refs), nscenes},
                          ctx->stream, d_rays, n, d_hits, d_segs, m, d_visible, GR_OCCLUSION_EXACT, &best);
    if (d_refs) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(d_refs);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "autotune");
    tuning_out(best, out);
    return GR_OK;
}

int gr_scene_autotune(gr_scene* s, const gr_ray* d_rays, uint32_t n, gr_hit* d_hits, gr_tuning* out) {
    if (s && n == 0) return fail(s->ctx, GR_ERR_INVALID, "autotune needs rays and a hit buffer%s");
    return gr_scene_autotune_mixed(s, d_rays, n, d_hits, nullptr, 0, nullptr, GR_OCCLUSION_EXACT, out);
}

int gr_film_create(gr_ctx* ctx, uint32_t width, uint32_t height, gr_film** out) {
    if (!ctx || !out) return GR_ERR_INVALID;
    *out = nullptr;
    if (ctx->device == GR_HOST_ONLY) return fail(ctx, GR_ERR_STATE, "host-only context%s");
    if (width == 0 || height == 0 || (uint64_t)width * height > 0x7fffffffull)
        return fail(ctx, GR_ERR_INVALID, "bad film size%s");
    gr_film* f = new (std::nothrow) gr_film();
    if (!f) return fail(ctx, GR_ERR_NOMEM, "out of host memory%s");
    f->ctx = ctx;
    f->width = width;
    f->height = height;
    const size_t bytes = (size_t)width * height * sizeof(float4);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&f->d_pixels), bytes);
    if (e == hipSuccess) e = hipMemsetAsync(f->d_pixels, 0, bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(f->d_pixels);
        delete f;
        return hip_fail(ctx, e, "film allocation");
    }
    gr::ctx_retain(ctx);
    *out = f;
    return GR_OK;
}

void gr_film_destroy(gr_film* f) {
    if (!f) return;
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    if (f->copy_stream) (void)hipStreamSynchronize(f->copy_stream);
    (void)hipFree(f->d_snapshot);
    (void)hipHostFree(f->h_pinned);
    if (f->snap_ready) (void)hipEventDestroy(f->snap_ready);
    if (f->copy_done) (void)hipEventDestroy(f->copy_done);
    if (f->copy_stream) (void)hipStreamDestroy(f->copy_stream);
    (void)hipFree(f->d_pixels);
    release_render_buffers(f->rb);
    gr_ctx* ctx = f->ctx;
    delete f;
    gr::ctx_release(ctx);
}

int gr_film_clear(gr_film* f) {
    if (!f) return GR_ERR_INVALID;
    hipError_t e = hipMemsetAsync(f->d_pixels, 0, (size_t)f->width * f->height * sizeof(float4), f->ctx->stream);
    return e == hipSuccess ? GR_OK : hip_fail(f->ctx, e, "film clear");
}

int gr_film_readback(gr_film* f, float* rgbw) {
    if (!f || !rgbw) return GR_ERR_INVALID;
    hipError_t e = hipMemcpyAsync(rgbw, f->d_pixels, (size_t)f->width * f->height * sizeof(float4),
                                  hipMemcpyDeviceToHost, f->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->ctx->stream);
    return e == hipSuccess ? GR_OK : hip_fail(f->ctx, e, "film readback");
}

int gr_film_readback_begin(gr_film* f) {
    if (!f) return GR_ERR_INVALID;
    if (f->copy_pending) return fail(f->ctx, GR_ERR_STATE, "a readback is already in flight%s");
    const size_t bytes = (size_t)f->width * f->height * sizeof(float4);
    hipError_t e = hipSuccess;
    if (!f->copy_stream) {
        e = hipSetDevice(f->ctx->device);
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&f->d_snapshot), bytes);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&f->h_pinned), bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&f->snap_ready, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&f->copy_done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&f->copy_stream, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(f->ctx, e, "readback buffers");
    }
    // the snapshot is ordered after the work queued so far; the host copy runs
    // on the side stream while the context's stream goes on rendering
    e = hipMemcpyAsync(f->d_snapshot, f->d_pixels, bytes, hipMemcpyDeviceToDevice, f->ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(f->snap_ready, f->ctx->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(f->copy_stream, f->snap_ready, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(f->h_pinned, f->d_snapshot, bytes, hipMemcpyDeviceToHost, f->copy_stream);
    if (e == hipSuccess) e = hipEventRecord(f->copy_done, f->copy_stream);
    if (e != hipSuccess) return hip_fail(f->ctx, e, "readback begin");
    f->copy_pending = true;
    return GR_OK;
}

int gr_film_readback_end(gr_film* f, float* rgbw) {
    if (!f || !rgbw) return GR_ERR_INVALID;
    if (!f->copy_pending) return fail(f->ctx, GR_ERR_STATE, "no readback in flight%s");
    f->copy_pending = false;
    const hipError_t e = hipEventSynchronize(f->copy_done);
    if (e != hipSuccess) return hip_fail(f->ctx, e, "readback end");
    std::memcpy(rgbw, f->h_pinned, (size_t)f->width * f->height * sizeof(float4));
    return GR_OK;
}

int gr_exr_write(const float* rgbw, uint32_t width, uint32_t height, const char* path) {
    if (!rgbw || !path || width == 0 || height == 0) return GR_ERR_INVALID;
    return write_exr(rgbw, width, height, path) ? GR_OK : GR_ERR_INVALID;
}

int gr_film_upload(gr_film* f, const float* rgbw) {
    if (!f || !rgbw) return GR_ERR_INVALID;
    hipError_t e = hipMemcpyAsync(f->d_pixels, rgbw, (size_t)f->width * f->height * sizeof(float4),
                                  hipMemcpyHostToDevice, f->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->ctx->stream);
    return e == hipSuccess ? GR_OK : hip_fail(f->ctx, e, "film upload");
}

int gr_trace(gr_scene* s, const gr_camera* cam, const gr_trace_params* prm, gr_film* film) {
    if (!s || !cam || !prm || !film) return GR_ERR_INVALID;
    gr_ctx* ctx = s->ctx;
    if (!s->committed) return fail(ctx, GR_ERR_STATE, "scene not committed%s");
    if (ctx->device == GR_HOST_ONLY) return fail(ctx, GR_ERR_STATE, "host-only context%s");
    if (film->ctx != ctx) return fail(ctx, GR_ERR_INVALID, "film belongs to another context%s");
    if (prm->integrator > GR_INTEGRATOR_BDPT_LT) return fail(ctx, GR_ERR_INVALID, "unknown integrator%s");
    if (prm->integrator == GR_INTEGRATOR_GUIDED && (!prm->guide || prm->guide->ctx != ctx))
        return fail(ctx, GR_ERR_INVALID, "guided tracing needs a guide of the same context%s");
    if (prm->integrator != GR_INTEGRATOR_PT && prm->integrator != GR_INTEGRATOR_DISPLAY && s->shade.num_lights == 0)
        return fail(ctx, GR_ERR_INVALID, "integrator samples lights but the scene has none%s");
    if (!(cam->focal_length > 0.0f)) return fail(ctx, GR_ERR_INVALID, "camera focal length must be positive%s");
    if (prm->spp == 0) return GR_OK;
    hipError_t e = render_trace(s, *cam, *prm, film);
    if (e == hipErrorNotSupported) return fail(ctx, GR_ERR_INVALID, "integrator not available%s");
    return e == hipSuccess ? GR_OK : hip_fail(ctx, e, "trace");
}

int gr_render_jobs(gr_ctx* ctx, const gr_job* jobs, uint32_t njobs, const gr_batch_params* prm, double* job_ms) {
    if (!ctx || !prm || (njobs && !jobs)) return GR_ERR_INVALID;
    if (ctx->device == GR_HOST_ONLY) return fail(ctx, GR_ERR_STATE, "host-only context%s");
    if (njobs == 0) return GR_OK;
    if (prm->integrator != GR_INTEGRATOR_PT && prm->integrator != GR_INTEGRATOR_PT_NEE)
        return fail(ctx, GR_ERR_INVALID, "batched jobs use GR_INTEGRATOR_PT or GR_INTEGRATOR_PT_NEE%s");
    if (prm->tuning.lds_stack != 0 && prm->tuning.lds_stack != 8 && prm->tuning.lds_stack != 16 &&
        prm->tuning.lds_stack != 32)
        return fail(ctx, GR_ERR_INVALID, "tuning out of range%s");
    if (prm->job_order != GR_JOBS_AS_GIVEN && prm->job_order != GR_JOBS_FEWEST_FIRST)
        return fail(ctx, GR_ERR_INVALID, "unknown job order%s");
    for (uint32_t j = 0; j < njobs; ++j) {
        const gr_job& jb = jobs[j];
        if (!jb.scene || !jb.film) return fail(ctx, GR_ERR_INVALID, "job without scene or film%s");
        if (jb.scene->ctx != ctx || jb.film->ctx != ctx)
            return fail(ctx, GR_ERR_INVALID, "job scene or film belongs to another context%s");
        if (!jb.scene->committed) return fail(ctx, GR_ERR_STATE, "scene not committed%s");
        if (prm->integrator == GR_INTEGRATOR_PT_NEE && jb.scene->shade.num_lights == 0)
            return fail(ctx, GR_ERR_INVALID, "integrator samples lights but a scene has none%s");
        if (!(jb.camera.focal_length > 0.0f)) return fail(ctx, GR_ERR_INVALID, "camera focal length must be positive%s");
        if ((uint64_t)jb.film->width * jb.film->height * jb.spp > 0xffffffffull)
            return fail(ctx, GR_ERR_INVALID, "a job has more than 2^32 samples%s");
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = render_jobs(ctx, jobs, njobs, *prm, job_ms);
    return e == hipSuccess ? GR_OK : hip_fail(ctx, e, "batched render");
}

int gr_film_write_exr(gr_film* f, const char* path) {
    if (!f || !path) return GR_ERR_INVALID;
    std::vector<float> px((size_t)f->width * f->height * 4);
    int r = gr_film_readback(f, px.data());
    if (r != GR_OK) return r;
    if (!write_exr(px.data(), f->width, f->height, path))
        return fail(f->ctx, GR_ERR_INVALID, "cannot write %s", path);
    return GR_OK;
}

int gr_material_query(gr_scene* s, uint32_t prim, const float* d_in, uint32_t n, float* d_out) {
    if (!s) return GR_ERR_INVALID;
    if (!s->committed) return fail(s->ctx, GR_ERR_STATE, "scene not committed%s");
    if (s->ctx->device == GR_HOST_ONLY) return fail(s->ctx, GR_ERR_STATE, "host-only context%s");
    if (prim >= s->prims.size()) return fail(s->ctx, GR_ERR_INVALID, "prim index out of range%s");
    if (n && (!d_in || !d_out)) return fail(s->ctx, GR_ERR_INVALID, "null buffer%s");
    hipError_t e = material_query(s, prim, d_in, n, d_out);
    return e == hipSuccess ? GR_OK : hip_fail(s->ctx, e, "material query");
}

int gr_scene_take_overflow_count(gr_scene* s, uint64_t* out) {
    if (!s || !out) return GR_ERR_INVALID;
    if (!s->committed) return fail(s->ctx, GR_ERR_STATE, "scene not committed%s");
    if (s->ctx->device == GR_HOST_ONLY) return fail(s->ctx, GR_ERR_STATE, "host-only context%s");
    unsigned long long v = 0;
    hipError_t e = hipMemcpyAsync(&v, s->d_overflow, sizeof(v), hipMemcpyDeviceToHost, s->ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->d_overflow, 0, sizeof(v), s->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "overflow readback");
    *out = v;
    return GR_OK;
}

}  // extern "C"
