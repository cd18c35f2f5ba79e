// gr_sort.hip -- coherence sort of ray batches before traversal.
//
// Rays (or shadow segments) are keyed by the octant of their direction and the
// Morton code of their origin quantised to the scene box, then sorted by a
// stable LSD radix sort written for 64-lane waves.  The traversal kernels read
// rays through the resulting permutation and write every result back to the
// ray's own index, so sorting changes only the order in which rays run, never
// a result.  The sort allocates nothing and issues no memset, so it can sit
// inside a captured graph; its scratch is reserved beforehand.
#pragma clang fp contract(off)

#include "gr_internal.h"

namespace gr {
namespace {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kRounds = 16;               // rounds of 64 keys per wave
constexpr int kSeg = 64 * kRounds;        // keys per wave segment
constexpr uint32_t kTile = kSeg * kWaves;  // keys per block

__device__ __forceinline__ uint32_t live_count(uint32_t n_max, const uint32_t* n_ptr) {
    if (!n_ptr) return n_max;
    const uint32_t n = __builtin_amdgcn_readfirstlane(*n_ptr);
    return n < n_max ? n : n_max;
}

// Spreads the low 10 bits of x to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t quantise(float v, float lo, float scale, uint32_t top) {
    float f = (v - lo) * scale;
    f = fmaxf(f, 0.0f);  // also maps NaN to 0
    f = fminf(f, (float)top);
    return (uint32_t)f;
}

// key = direction octant (3 bits) above the origin's Morton code (3 * m bits).
__global__ __launch_bounds__(256) void k_sort_keys(const float4* __restrict__ in, uint32_t n_max,
                                                   const uint32_t* __restrict__ n_ptr, uint32_t segments,
                                                   float4 lo, float4 scale, uint32_t m,
                                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t n = live_count(n_max, n_ptr);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = in[2 * (size_t)i], b = in[2 * (size_t)i + 1];
    float dx = b.x, dy = b.y, dz = b.z;
    if (segments) {
        dx = b.x - a.x;
        dy = b.y - a.y;
        dz = b.z - a.z;
    }
    const uint32_t oct = (dx < 0.0f ? 4u : 0u) | (dy < 0.0f ? 2u : 0u) | (dz < 0.0f ? 1u : 0u);
    const uint32_t top = (1u << m) - 1u;
    const uint32_t qx = quantise(a.x, lo.x, scale.x, top);
    const uint32_t qy = quantise(a.y, lo.y, scale.y, top);
    const uint32_t qz = quantise(a.z, lo.z, scale.z, top);
    keys[i] = (oct << (3 * m)) | (spread3(qx) << 2) | (spread3(qy) << 1) | spread3(qz);
    vals[i] = i;
}

// Multi-scene keys: scene index (sbits) above octant and Morton code, the
// origin quantised to its own scene's box.
__global__ __launch_bounds__(256) void k_sort_keys_multi(const float4* __restrict__ in, uint32_t n_max,
                                                         const uint32_t* __restrict__ n_ptr, uint32_t segments,
                                                         const SceneRef* __restrict__ scenes, uint32_t nscenes,
                                                         uint32_t m, uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ vals) {
    const uint32_t n = live_count(n_max, n_ptr);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = in[2 * (size_t)i], b = in[2 * (size_t)i + 1];
    float dx = b.x, dy = b.y, dz = b.z;
    if (segments) {
        dx = b.x - a.x;
        dy = b.y - a.y;
        dz = b.z - a.z;
    }
    uint32_t si = __float_as_uint(a.w);
    si = si < nscenes ? si : 0u;
    const SceneRef& sr = scenes[si];
    const uint32_t oct = (dx < 0.0f ? 4u : 0u) | (dy < 0.0f ? 2u : 0u) | (dz < 0.0f ? 1u : 0u);
    const uint32_t top = (1u << m) - 1u;
    uint32_t q[3];
    const float o[3] = {a.x, a.y, a.z};
    for (int k = 0; k < 3; ++k) {
        const float ext = sr.root_hi[k] - sr.root_lo[k];
        q[k] = quantise(o[k], sr.root_lo[k], ext > 0.0f ? (float)(1u << m) / ext : 0.0f, top);
    }
    keys[i] = (si << (3 + 3 * m)) | (oct << (3 * m)) | (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    vals[i] = i;
}

// Per-block digit histogram, stored digit-major: hist[d * nblocks + block].
__global__ __launch_bounds__(kThreads) void k_sort_hist(const uint32_t* __restrict__ keys, uint32_t n_max,
                                                        const uint32_t* __restrict__ n_ptr, uint32_t shift,
                                                        uint32_t mask, uint32_t* __restrict__ hist,
                                                        uint32_t nblocks) {
    __shared__ uint32_t h[256];
    const uint32_t n = live_count(n_max, n_ptr);
    h[threadIdx.x] = 0u;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
    for (uint32_t i = 0; i < kTile / kThreads; ++i) {
        const size_t idx = base + (size_t)i * kThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & mask], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Inclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kThreads; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    const uint32_t r = s[threadIdx.x];
    __syncthreads();
    return r;
}

// One block per digit: exclusive scan of that digit's row of block counts;
// the row total goes to totals[d].
__global__ __launch_bounds__(kThreads) void k_sort_scan(uint32_t* __restrict__ hist, uint32_t nblocks,
                                                        uint32_t* __restrict__ totals) {
    __shared__ uint32_t s[kThreads];
    __shared__ uint32_t s_total;
    uint32_t* row = hist + (size_t)blockIdx.x * nblocks;
    uint32_t carry = 0;
    for (uint32_t c = 0; c < nblocks; c += kThreads) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t v = i < nblocks ? row[i] : 0u;
        const uint32_t incl = block_scan(v, s);
        if (threadIdx.x == kThreads - 1) s_total = incl;
        if (i < nblocks) row[i] = carry + incl - v;
        __syncthreads();
        carry += s_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Stable scatter.  Each wave ranks its 1024-key segment round by round: lanes
// holding the same digit find each other with one ballot per digit bit, and
// the wave's running count per digit lives in LDS.  Keys keep their order
// because blocks, wave segments, rounds and lanes cover the input in index
// order.
__global__ __launch_bounds__(kThreads) void k_sort_scatter(const uint32_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                           uint32_t n_max, const uint32_t* __restrict__ n_ptr,
                                                           uint32_t shift, uint32_t mask, uint32_t dbits,
                                                           const uint32_t* __restrict__ hist,
                                                           const uint32_t* __restrict__ totals, uint32_t nblocks) {
    __shared__ uint32_t s_scan[kThreads];
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_cnt[kWaves][256];
    const uint32_t n = live_count(n_max, n_ptr);
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    {
        const uint32_t v = totals[tid];
        const uint32_t incl = block_scan(v, s_scan);
        s_base[tid] = incl - v + hist[(size_t)tid * nblocks + blockIdx.x];
        for (int ww = 0; ww < kWaves; ++ww) s_cnt[ww][tid] = 0u;
    }
    __syncthreads();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const size_t seg0 = (size_t)blockIdx.x * kTile + (size_t)w * kSeg;
    uint32_t key[kRounds], val[kRounds], rank[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const size_t idx = seg0 + (size_t)r * 64 + lane;
        const bool valid = idx < n;
        const uint32_t k = valid ? kin[idx] : 0u;
        const uint32_t v = valid ? vin[idx] : 0u;
        const uint32_t d = (k >> shift) & mask;
        uint64_t peers = __ballot(valid);
        for (uint32_t b = 0; b < dbits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t within = (uint32_t)__popcll(peers & below);
        // Every lane of the wave reads the running count before the lowest lane
        // of each digit group writes it: LDS serves one wave's accesses in
        // program order.
        const uint32_t before = valid ? s_cnt[w][d] : 0u;
        if (valid && within == 0) s_cnt[w][d] = before + (uint32_t)__popcll(peers);
        key[r] = k;
        val[r] = v;
        rank[r] = before + within;
    }
    __syncthreads();
    {
        uint32_t acc = 0;
        for (int ww = 0; ww < kWaves; ++ww) {
            const uint32_t c = s_cnt[ww][tid];
            s_cnt[ww][tid] = acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const size_t idx = seg0 + (size_t)r * 64 + lane;
        if (idx < n) {
            const uint32_t d = (key[r] >> shift) & mask;
            const uint32_t pos = s_base[d] + s_cnt[w][d] + rank[r];
            if (pos < n) {  // always true for consistent counts; keeps a fault out of reach
                kout[pos] = key[r];
                vout[pos] = val[r];
            }
        }
    }
}

inline uint32_t ceil_div(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

uint32_t sort_key_bits(uint32_t bits) {
    if (bits < 6) return 0;
    if (bits > 30) bits = 30;
    return 3 + 3 * ((bits - 3) / 3);
}

size_t sort_scratch_bytes(size_t n) {
    const size_t nblocks = (n + kTile - 1) / kTile;
    const size_t keys = ((n * 4 + 255) / 256) * 256;
    const size_t hist = ((nblocks * 256 * 4 + 255) / 256) * 256;
    return 4 * keys + hist + 256 * 4;
}

void carve_sort(SortScratch& ss, void* base, size_t n) {
    const size_t nblocks = (n + kTile - 1) / kTile;
    const size_t keys = ((n * 4 + 255) / 256) * 256;
    const size_t hist = ((nblocks * 256 * 4 + 255) / 256) * 256;
    char* p = static_cast<char*>(base);
    ss.keys[0] = reinterpret_cast<uint32_t*>(p);
    ss.keys[1] = reinterpret_cast<uint32_t*>(p + keys);
    ss.vals[0] = reinterpret_cast<uint32_t*>(p + 2 * keys);
    ss.vals[1] = reinterpret_cast<uint32_t*>(p + 3 * keys);
    ss.hist = reinterpret_cast<uint32_t*>(p + 4 * keys);
    ss.totals = reinterpret_cast<uint32_t*>(p + 4 * keys + hist);
    ss.cap = n;
}

hipError_t reserve_sort(SortScratch& ss, size_t n) {
    if (n <= ss.cap) return hipSuccess;
    release_sort(ss);
    char* p = nullptr;
    hipError_t e = hipMalloc(&p, sort_scratch_bytes(n));
    if (e != hipSuccess) return e;
    ss.blob = p;
    carve_sort(ss, p, n);
    return hipSuccess;
}

void release_sort(SortScratch& ss) {
    (void)hipFree(ss.blob);
    ss = SortScratch{};
}

namespace {
// Radix passes over the first `bits` key bits of ss.keys[0] / ss.vals[0].
hipError_t radix_passes(uint32_t n_max, const uint32_t* n_ptr, uint32_t bits, SortScratch& ss, const uint32_t** perm,
                        hipStream_t s) {
    const uint32_t nblocks = ceil_div(n_max, kTile);
    int src = 0;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        const uint32_t dbits = bits - shift < 8 ? bits - shift : 8;
        const uint32_t mask = (1u << dbits) - 1u;
        hipLaunchKernelGGL(k_sort_hist, dim3(nblocks), dim3(kThreads), 0, s, ss.keys[src], n_max, n_ptr, shift,
                           mask, ss.hist, nblocks);
        hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(kThreads), 0, s, ss.hist, nblocks, ss.totals);
        hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(kThreads), 0, s, ss.keys[src], ss.vals[src],
                           ss.keys[src ^ 1], ss.vals[src ^ 1], n_max, n_ptr, shift, mask, dbits, ss.hist,
                           ss.totals, nblocks);
        src ^= 1;
    }
    *perm = ss.vals[src];
    return hipGetLastError();
}
}  // namespace

hipError_t sort_rays_multi(const SceneRef* scenes, uint32_t nscenes, const float4* in, uint32_t n_max,
                           const uint32_t* n_ptr, bool segments, uint32_t bits, SortScratch& ss,
                           const uint32_t** perm, hipStream_t s) {
    bits = sort_key_bits(bits);
    *perm = nullptr;
    if (bits == 0 || n_max == 0) return hipSuccess;
    if (ss.cap < n_max) return hipErrorInvalidValue;
    uint32_t sbits = 0;
    while ((1u << sbits) < nscenes) ++sbits;
    uint32_t m = (bits - 3) / 3;
    while (m > 0 && sbits + 3 + 3 * m > 32) --m;
    hipLaunchKernelGGL(k_sort_keys_multi, dim3(ceil_div(n_max, 256)), dim3(256), 0, s, in, n_max, n_ptr,
                       segments ? 1u : 0u, scenes, nscenes, m, ss.keys[0], ss.vals[0]);
    return radix_passes(n_max, n_ptr, sbits + 3 + 3 * m, ss, perm, s);
}

hipError_t sort_rays(const SceneView& sv, const float4* in, uint32_t n_max, const uint32_t* n_ptr, bool segments,
                     uint32_t bits, SortScratch& ss, const uint32_t** perm, hipStream_t s) {
    bits = sort_key_bits(bits);
    *perm = nullptr;
    if (bits == 0 || n_max == 0) return hipSuccess;
    if (ss.cap < n_max) return hipErrorInvalidValue;
    const uint32_t m = (bits - 3) / 3;
    float4 lo, scale;
    float* l = &lo.x;
    float* sc = &scale.x;
    for (int k = 0; k < 3; ++k) {
        const float ext = sv.root_hi[k] - sv.root_lo[k];
        l[k] = sv.root_lo[k];
        sc[k] = ext > 0.0f ? (float)(1u << m) / ext : 0.0f;
    }
    lo.w = scale.w = 0.0f;
    hipLaunchKernelGGL(k_sort_keys, dim3(ceil_div(n_max, 256)), dim3(256), 0, s, in, n_max, n_ptr,
                       segments ? 1u : 0u, lo, scale, m, ss.keys[0], ss.vals[0]);
    return radix_passes(n_max, n_ptr, bits, ss, perm, s);
}

}  // namespace gr
