// gr_noalloc.hip -- closest-hit traversal that allocates nothing.
//
// gr_intersect_closest keeps its ray-sort scratch (keys, values, digit
// histograms) in a per-scene buffer it grows with hipMalloc, and the
// persistent traversal's work counters in a per-scene allocation.  Callers that
// cannot allocate -- code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc -- use gr_intersect_closest_noalloc instead:
// the same kernels, with all of that scratch in one file-scope device array.
//
// The array's host-side name is not a device address.  gr_ctx_create calls
// noalloc_setup() once per device, with that device current; it resolves the
// array's address on that device with hipGetSymbolAddress and keeps it.  Only
// that address is ever handed to kernels.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two host threads) race.  The work counters are cleared by a
// kernel at the start of each launch; the sort writes every key and value it
// reads, so no other content needs clearing.
#pragma clang fp contract(off)

#include <mutex>

#include "gr_internal.h"

namespace {

// Scratch layout: the traversal's work counters, then the sort scratch for
// GR_NOALLOC_MAX_RAYS keys (4 key/value arrays, digit histograms, totals).
constexpr size_t kCounterBytes = 1024;
constexpr size_t kSortBytes = 16u * GR_NOALLOC_MAX_RAYS + 256u * 4u * (GR_NOALLOC_MAX_RAYS / 4096u) + 1024u;
constexpr size_t kScratchBytes = kCounterBytes + kSortBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

int fail(gr_ctx* ctx, int code, const char* msg) {
    if (ctx) ctx->err = msg;
    return code;
}

}  // namespace

namespace gr {

hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return hipSuccess;
    if (sort_scratch_bytes(GR_NOALLOC_MAX_RAYS) > kSortBytes) return hipErrorInvalidValue;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_ptr[device] = p;
    return hipSuccess;
}

}  // namespace gr

extern "C" int gr_intersect_closest_noalloc(gr_scene* s, const gr_ray* d_rays, uint32_t n, gr_hit* d_hits,
                                            uint32_t* d_counts) {
    if (!s) return GR_ERR_INVALID;
    if (n > GR_NOALLOC_MAX_RAYS) return fail(s->ctx, GR_ERR_INVALID, "more rays than GR_NOALLOC_MAX_RAYS");
    if (!s->committed) return fail(s->ctx, GR_ERR_STATE, "scene not committed");
    if (s->ctx->device == GR_HOST_ONLY) return fail(s->ctx, GR_ERR_STATE, "host-only context");
    if (n && (!d_rays || !d_hits)) return fail(s->ctx, GR_ERR_INVALID, "null ray/hit buffer");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (s->ctx->device >= 0 && s->ctx->device < kMaxDevices) base = g_ptr[s->ctx->device];
    }
    if (!base) return fail(s->ctx, GR_ERR_STATE, "device scratch not set up for this context's device");
    char* p = static_cast<char*>(base);
    gr::SortScratch ss{};
    gr::carve_sort(ss, p + kCounterBytes, GR_NOALLOC_MAX_RAYS);
    gr::TraceSetup ts{s->view, s->tuning, &ss, nullptr};
    ts.sv.next_ray = reinterpret_cast<uint32_t*>(p);
    const hipError_t e = gr::launch_intersect(ts, d_rays, n, nullptr, d_hits, d_counts, s->ctx->stream);
    if (e != hipSuccess) {
        s->ctx->err = std::string("intersect launch: ") + hipGetErrorString(e);
        return GR_ERR_HIP;
    }
    return GR_OK;
}

static_assert(gr::kRayCounterBytes <= kCounterBytes, "counter region");
