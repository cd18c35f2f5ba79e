// gr_guide.hip -- the learned path guide on gfx950: inference, training
// (forward, backward, weight gradients) and Adam, all on the matrix cores
// (fp16 operands, fp32 accumulation: v_mfma_f32_16x16x32_f16 for inference,
// v_mfma_f32_16x16x16_f16 for training).
//
// Layout: one wave evaluates 16-query tiles (inference: two at a time, so each
// weight fragment read from LDS feeds two MFMAs).  Activations are held
// transposed (rows = features, columns = queries), so the C tiles a layer
// produces are, converted to fp16, exactly the B tiles the next layer consumes
// -- no data movement between layers.  Weights are stored pre-swizzled as A
// fragments: W^T tiles in K=32 order for inference, W^T tiles (64 lanes x 4
// halves per 16x16 tile) for the training forward pass, W tiles for
// back-propagation.  Weight gradients are a GEMM over the
// batch dimension (dW_l = A_{l-1}^T dZ_l); the training kernel writes both
// operands transposed into fragment order through a per-wave LDS tile and a
// split-K kernel reduces them.
#pragma clang fp contract(off)

#include <cstring>
#include <new>
#include <vector>

#include "gr_guide.h"
#include "gr_internal.h"

namespace gr {
namespace {

constexpr int kInferThreads = 768;  // 12 waves sharing one LDS copy of the inference weights
constexpr int kTrainThreads = 256;
constexpr int kAFrags = 29;  // layer-input fragments per 16-query tile: 13 + 4 * 4
constexpr int kZFrags = 18;  // dZ fragments per tile: 4 * 4 + 2
constexpr int kDwTiles = 108;
constexpr int kDwSplit = 16;
constexpr int kGuideBiases = 4 * kGuideWidth + kGuideOut;  // 288, in the forward pass's order
constexpr float kGradClip = 1.0e4f;
constexpr float kValueClip = 20.0f;

__device__ __forceinline__ float4_t mfma16(half4_t a, half4_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float4_t mfma32(half8_t a, half8_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ half8_t cat8(half4_t lo, half4_t hi) {
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ __forceinline__ float4_t zero4() {
    float4_t z;
    z[0] = z[1] = z[2] = z[3] = 0.0f;
    return z;
}

// Orders one wave's LDS writes before its later reads of other lanes' data.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Biases gathered for the forward pass: b1..b4 (64 each) then b5 (32).
__device__ __forceinline__ float bias_at(const float* params, int i) {
    return i < 4 * kGuideWidth ? params[guide_b_off(i / kGuideWidth) + i % kGuideWidth]
                               : params[kB5 + (i - 4 * kGuideWidth)];
}

// Biases indexed as the forward pass wants them, read from the parameters.
struct BiasView {
    const float* p;
    __device__ __forceinline__ float operator[](int k) const { return bias_at(p, k); }
};

// Forward pass of one 16-query tile.  W: forward fragments (LDS or global),
// bias: 288 biases, x: this lane's query's raw inputs.  Keeps the hidden
// activations (B layout) for back-propagation; the encoded inputs are
// recomputed where needed.  o: the 32 outputs (C layout).
template <typename WP, typename BP>
__device__ __forceinline__ void guide_forward(WP W, BP bias, const _Float16* ftile, uint32_t lane, half4_t h[4][4],
                                              float4_t o[2]) {
    const uint32_t g = lane >> 4;
    const _Float16* frow = ftile + (lane & 15u) * kFeatStride + 4 * g;
    float4_t acc[4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) acc[ot] = zero4();
#pragma unroll
    for (int kb = 0; kb < 13; ++kb) {
        const half4_t f = *reinterpret_cast<const half4_t*>(frow + kb * 16);
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) acc[ot] = mfma16(W[(guide_fwd_base(0) + ot * 13 + kb) * 64 + lane], f, acc[ot]);
    }
#pragma unroll
    for (int ot = 0; ot < 4; ++ot)
#pragma unroll
        for (int c = 0; c < 4; ++c) h[0][ot][c] = (_Float16)fmaxf(acc[ot][c] + bias[ot * 16 + 4 * g + c], 0.0f);
#pragma unroll
    for (int l = 1; l < 4; ++l) {
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) acc[ot] = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
#pragma unroll
            for (int ot = 0; ot < 4; ++ot)
                acc[ot] = mfma16(W[(guide_fwd_base(l) + ot * 4 + kb) * 64 + lane], h[l - 1][kb], acc[ot]);
#pragma unroll
        for (int ot = 0; ot < 4; ++ot)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                h[l][ot][c] = (_Float16)fmaxf(acc[ot][c] + bias[l * kGuideWidth + ot * 16 + 4 * g + c], 0.0f);
    }
    o[0] = zero4();
    o[1] = zero4();
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
        for (int ot = 0; ot < 2; ++ot) o[ot] = mfma16(W[(guide_fwd_base(4) + ot * 4 + kb) * 64 + lane], h[3][kb], o[ot]);
#pragma unroll
    for (int ot = 0; ot < 2; ++ot)
#pragma unroll
        for (int c = 0; c < 4; ++c) o[ot][c] = o[ot][c] + bias[4 * kGuideWidth + ot * 16 + 4 * g + c];
}

// The forward pass with the K=32 instruction (inference) for T 16-query
// tiles at once, so that each weight fragment read from LDS feeds T MFMAs.
// x0[t], x1[t]: raw inputs 2G and 2G+1 (G = lane / 16) of the lane's query in
// tile t; layer 1's operands are encoded in registers just before use.
template <int T>
__device__ __forceinline__ void guide_forward32(const half8_t* W, const float* bias, const float* x0,
                                                const float* x1, uint32_t lane, float4_t o[][2]) {
    const uint32_t g = lane >> 4;
    // each accumulator starts from its bias: rows 4g..4g+3 of output tile ot
    const float4_t* B4 = reinterpret_cast<const float4_t*>(bias);
    float4_t acc[T][4];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot) {
        const float4_t b = B4[ot * 4 + g];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t][ot] = b;
    }
#pragma unroll
    for (int kb = 0; kb < 7; ++kb) {
        half8_t f[T];
#pragma unroll
        for (int t = 0; t < T; ++t) f[t] = guide_encode_block(x0[t], x1[t], kb);
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) {
            const half8_t w = W[(guide_fwd32_base(0) + ot * 7 + kb) * 64 + lane];
#pragma unroll
            for (int t = 0; t < T; ++t) acc[t][ot] = mfma32(w, f[t], acc[t][ot]);
        }
    }
    half4_t h[T][4];
    const half4_t hz = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
#pragma unroll
    for (int l = 1; l <= 4; ++l) {
        // ReLU after the fp16 rounding: the same values as before it, in packed halves
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int ot = 0; ot < 4; ++ot)
                h[t][ot] = __builtin_elementwise_max(__builtin_convertvector(acc[t][ot], half4_t), hz);
        const int nout = l == 4 ? 2 : 4;
#pragma unroll
        for (int ot = 0; ot < nout; ++ot) {
            const float4_t b = B4[l * 16 + ot * 4 + g];
#pragma unroll
            for (int t = 0; t < T; ++t) acc[t][ot] = b;
        }
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
            for (int ot = 0; ot < nout; ++ot) {
                const half8_t w = W[(guide_fwd32_base(l) + ot * 2 + kb) * 64 + lane];
#pragma unroll
                for (int t = 0; t < T; ++t) acc[t][ot] = mfma32(w, cat8(h[t][2 * kb], h[t][2 * kb + 1]), acc[t][ot]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int ot = 0; ot < 2; ++ot) o[t][ot] = acc[t][ot];
}

// One lobe's parameters from its 4 outputs (logit, log kappa, theta, phi).
struct Lobe {
    float logit, lk_raw, kappa, sa, sb, theta, phi;
    vec3 mu;
};

__device__ __forceinline__ Lobe head_lobe(float4_t o) {
    Lobe L;
    L.logit = o[0];
    L.lk_raw = o[1];
    L.kappa = expf(fminf(fmaxf(o[1], kLogKappaLo), kLogKappaHi));
    L.sa = guide_sigmoid(o[2]);
    L.sb = guide_sigmoid(o[3]);
    L.theta = kPi * L.sa;
    L.phi = 2.0f * kPi * L.sb;
    L.mu = guide_mu(L.theta, L.phi);
    return L;
}

// Softmax weights of the lane's two lobes over the 8 lobes of its query (the
// query's four lanes are lane, lane ^ 16, lane ^ 32, lane ^ 48).
__device__ __forceinline__ void head_weights(const Lobe& a, const Lobe& b, float* wa, float* wb) {
    float m = fmaxf(a.logit, b.logit);
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    const float ea = expf(a.logit - m), eb = expf(b.logit - m);
    float s = ea + eb;
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    *wa = ea / s;
    *wb = eb / s;
}

// The inference head on the hardware exp2, reciprocal, sin and cos (angles in
// revolutions): theta = pi sa is half a revolution times sa, phi = 2 pi sb is
// sb revolutions.  The sampler and the pdf both read the mixture it writes,
// so their approximation errors (a few ulp) cannot bias a guided estimate.
__device__ __forceinline__ float hw_exp(float v) { return __builtin_amdgcn_exp2f(v * 1.44269504088896341f); }
__device__ __forceinline__ float hw_sigmoid(float v) { return __builtin_amdgcn_rcpf(1.0f + hw_exp(-v)); }

struct FastLobe {
    float logit, kappa;
    vec3 mu;
};

__device__ __forceinline__ FastLobe head_lobe_fast(float4_t o) {
    FastLobe L;
    L.logit = o[0];
    L.kappa = hw_exp(fminf(fmaxf(o[1], kLogKappaLo), kLogKappaHi));
    const float ht = 0.5f * hw_sigmoid(o[2]), sb = hw_sigmoid(o[3]);
    const float st = __builtin_amdgcn_sinf(ht), sp = __builtin_amdgcn_sinf(sb), cp = __builtin_amdgcn_cosf(sb);
    L.mu = mk(st * cp, st * sp, __builtin_amdgcn_cosf(ht));
    return L;
}

__device__ __forceinline__ void head_weights_fast(const FastLobe& a, const FastLobe& b, float* wa, float* wb) {
    float m = fmaxf(a.logit, b.logit);
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    const float ea = hw_exp(a.logit - m), eb = hw_exp(b.logit - m);
    float s = ea + eb;
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    const float r = __builtin_amdgcn_rcpf(s);
    *wa = ea * r;
    *wb = eb * r;
}

// Tiles of 16 queries per wave and pass: each weight fragment feeds that many MFMAs.
constexpr int kInferTiles = 2;

// Raw inputs 2G and 2G+1 of query q (G = lane / 16); zeros past the end.
__device__ __forceinline__ float2 load_pair(const float* raw, size_t q, size_t n, uint32_t g) {
    return q < n ? reinterpret_cast<const float2*>(raw)[q * 4 + g] : make_float2(0.0f, 0.0f);
}

__global__ __launch_bounds__(kInferThreads) void k_guide_infer(const uint16_t* __restrict__ fwd32,
                                                               const float* __restrict__ params,
                                                               const float* __restrict__ raw, uint32_t n_max,
                                                               const uint32_t* __restrict__ n_ptr,
                                                               float* __restrict__ mix) {
    __shared__ half8_t s_w[kFwd32Tiles * 64];
    __shared__ __attribute__((aligned(16))) float s_b[4 * kGuideWidth + kGuideOut];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < (uint32_t)kFwd32Tiles * 64; i += kInferThreads)
        s_w[i] = reinterpret_cast<const half8_t*>(fwd32)[i];
    for (uint32_t i = tid; i < (uint32_t)(4 * kGuideWidth + kGuideOut); i += kInferThreads)
        s_b[i] = bias_at(params, (int)i);
    __syncthreads();
    size_t n = n_max;
    if (n_ptr) {
        const uint32_t m = *n_ptr;
        n = m < n_max ? m : n_max;
    }
    const uint32_t lane = tid & 63u, g = lane >> 4;
    const size_t span = 16 * kInferTiles;
    const size_t waves = (size_t)gridDim.x * (kInferThreads / 64);
    for (size_t base = ((size_t)blockIdx.x * (kInferThreads / 64) + (tid >> 6)) * span; base < n;
         base += waves * span) {
        float x0[kInferTiles], x1[kInferTiles];
#pragma unroll
        for (int t = 0; t < kInferTiles; ++t) {
            const float2 v = load_pair(raw, base + 16 * t + (lane & 15u), n, g);
            x0[t] = v.x;
            x1[t] = v.y;
        }
        // the weight fragments stay in LDS: hiding the lane's offset from the
        // optimiser keeps it from hoisting all 56 reads out of the loop (and
        // spilling them to scratch)
        uint32_t wl = lane;
        asm volatile("" : "+v"(wl));
        float4_t o[kInferTiles][2];
        guide_forward32<kInferTiles>(s_w, s_b, x0, x1, wl, o);
#pragma unroll
        for (int t = 0; t < kInferTiles; ++t) {
            const FastLobe la = head_lobe_fast(o[t][0]), lb = head_lobe_fast(o[t][1]);
            float wa, wb;
            head_weights_fast(la, lb, &wa, &wb);
            const size_t q = base + 16 * t + (lane & 15u);
            if (q < n) {
                // the lane's two lobes as one 40-byte run (lobe order within a mixture
                // is immaterial: slots 2g, 2g + 1 hold network lobes g, g + 4)
                float2* p2 = reinterpret_cast<float2*>(mix + q * kMixFloats + 10 * g);
                p2[0] = make_float2(wa, la.kappa);
                p2[1] = make_float2(la.mu.x, la.mu.y);
                p2[2] = make_float2(la.mu.z, wb);
                p2[3] = make_float2(lb.kappa, lb.mu.x);
                p2[4] = make_float2(lb.mu.y, lb.mu.z);
            }
        }
    }
}

// Writes a 16x16 tile held in C/B layout (lane: column lane%16, rows
// 4(lane/16)..+3) as a fragment whose lane holds row lane%16 and k =
// 4(lane/16)..+3 of the transposed tile, through the wave's LDS tile.
__device__ __forceinline__ void store_transposed(half4_t v, _Float16* lds, uint16_t* dst, uint32_t lane) {
    const uint32_t col = lane & 15u, g = lane >> 4;
#pragma unroll
    for (int c = 0; c < 4; ++c) lds[col * 16 + 4 * g + c] = v[c];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    half4_t r;
#pragma unroll
    for (int c = 0; c < 4; ++c) r[c] = lds[(4 * g + c) * 16 + col];
    reinterpret_cast<half4_t*>(dst)[lane] = r;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ half4_t to_half4(float4_t v) {
    half4_t h;
#pragma unroll
    for (int c = 0; c < 4; ++c) h[c] = (_Float16)v[c];
    return h;
}

// Sum over the 16 queries of a C-layout tile, stored to the tile's bias
// partials at [off + 4(lane/16) + c] (biases in the forward pass's order).
__device__ __forceinline__ void add_bias_grad(float4_t v, float* tile_bias, int off, uint32_t lane) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float s = v[c];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        s += __shfl_xor(s, 8);
        if ((lane & 15u) == 0) tile_bias[off + 4 * (lane >> 4) + c] = s;
    }
}

// Parameter index of bias i in the forward pass's order (b1..b4, then b5).
__host__ __device__ __forceinline__ int bias_param(int i) {
    return i < 4 * kGuideWidth ? guide_b_off(i / kGuideWidth) + i % kGuideWidth : kB5 + (i - 4 * kGuideWidth);
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Forward + backward of one 16-query tile per wave.  Writes the dW operands
// (layer inputs and dZ, transposed into fragment order) and the tile's bias
// gradients.  Records: the first n of `recs`; batch entry i uses record i
// (sequential) or a hashed draw (training).
__global__ __launch_bounds__(kTrainThreads) void k_guide_train(const uint16_t* __restrict__ fwd,
                                                               const uint16_t* __restrict__ bwd,
                                                               const float* __restrict__ params,
                                                               float* __restrict__ bias_part,
                                                               const float* __restrict__ recs, uint32_t n_max,
                                                               const uint32_t* __restrict__ n_ptr,
                                                               const float* __restrict__ stats, uint32_t batch,
                                                               uint32_t draw_key, uint32_t sequential,
                                                               uint16_t* __restrict__ at_buf,
                                                               uint16_t* __restrict__ dz_buf) {
    __shared__ _Float16 s_t[kTrainThreads / 64][256];
    __shared__ __attribute__((aligned(16))) _Float16 s_f[kTrainThreads / 64][16 * kFeatStride];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6, g = lane >> 4;
    const uint32_t tile = blockIdx.x * (kTrainThreads / 64) + wave;
    if (tile * 16 >= batch) return;
    uint32_t n = n_max;
    if (n_ptr) {
        const uint32_t m = *n_ptr;
        n = m < n_max ? m : n_max;
    }
    const uint32_t i = tile * 16 + (lane & 15u);
    bool valid = i < batch && n > 0;
    uint32_t ri = 0;
    if (valid) ri = sequential ? i : hash32(draw_key ^ hash32(i)) % n;
    if (sequential && ri >= n) valid = false;
    const float* rec = recs + (size_t)ri * kRecordFloats;
    float4 xa = make_float4(0.0f, 0.0f, 0.0f, 0.0f), xb = xa;
    if (valid) {
        xa = make_float4(rec[0], rec[1], rec[2], rec[3]);
        xb = make_float4(rec[4], rec[5], rec[6], rec[7]);
    }
    const vec3 dir = valid ? mk(rec[8], rec[9], rec[10]) : mk(0.0f, 0.0f, 1.0f);
    float vnorm = 0.0f;
    if (valid && stats[0] > 0.0f) vnorm = fminf(fmaxf(rec[11] * (float)n / stats[0], 0.0f), kValueClip);
    const half4_t* W = reinterpret_cast<const half4_t*>(fwd);
    const half4_t* Wb = reinterpret_cast<const half4_t*>(bwd);
    _Float16* ft = s_f[wave];
    guide_encode_tile(xa, xb, ft, lane);
    wave_lds_sync();
    half4_t h[4][4];
    float4_t o[2];
    guide_forward(W, BiasView{params}, ft, lane, h, o);

    // ---- output head: d(-vnorm log q)/d outputs
    const Lobe la = head_lobe(o[0]), lb = head_lobe(o[1]);
    float wa, wb;
    head_weights(la, lb, &wa, &wb);
    const float pa = wa * vmf_pdf(la.kappa, la.mu, dir), pb = wb * vmf_pdf(lb.kappa, lb.mu, dir);
    float q = pa + pb;
    q += __shfl_xor(q, 16);
    q += __shfl_xor(q, 32);
    const float s = (q > 1.0e-30f) ? -vnorm : 0.0f;
    const float ra = q > 1.0e-30f ? pa / q : 0.0f, rb = q > 1.0e-30f ? pb / q : 0.0f;
    float4_t dz5[2];
    {
        const Lobe* L2[2] = {&la, &lb};
        const float r2[2] = {ra, rb}, w2[2] = {wa, wb};
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const Lobe& L = *L2[j];
            const float r = r2[j];
            const float md = dot(L.mu, dir) - 1.0f;
            float dk = 0.0f;
            if (L.lk_raw > kLogKappaLo && L.lk_raw < kLogKappaHi)
                dk = s * r * (1.0f - 2.0f * L.kappa / expm1f(2.0f * L.kappa) + L.kappa * md);
            const float ct = cosf(L.theta), st = sinf(L.theta), cp = cosf(L.phi), sp = sinf(L.phi);
            const vec3 dmu_t = mk(ct * cp, ct * sp, -st);
            const vec3 dmu_p = mk(-st * sp, st * cp, 0.0f);
            const float dt = s * r * L.kappa * dot(dir, dmu_t) * kPi * L.sa * (1.0f - L.sa);
            const float dp = s * r * L.kappa * dot(dir, dmu_p) * 2.0f * kPi * L.sb * (1.0f - L.sb);
            dz5[j][0] = fminf(fmaxf(s * (r - w2[j]), -kGradClip), kGradClip);
            dz5[j][1] = fminf(fmaxf(dk, -kGradClip), kGradClip);
            dz5[j][2] = fminf(fmaxf(dt, -kGradClip), kGradClip);
            dz5[j][3] = fminf(fmaxf(dp, -kGradClip), kGradClip);
        }
    }

    uint16_t* at = at_buf + (size_t)tile * kAFrags * kFragHalves;
    uint16_t* dz = dz_buf + (size_t)tile * kZFrags * kFragHalves;
    float* tb = bias_part + (size_t)tile * kGuideBiases;
    _Float16* lds = s_t[wave];

    // layer 5 (index 4): dZ = dz5
    half4_t zh[4];
#pragma unroll
    for (int ob = 0; ob < 2; ++ob) {
        zh[ob] = to_half4(dz5[ob]);
        store_transposed(zh[ob], lds, dz + (16 + ob) * kFragHalves, lane);
        add_bias_grad(dz5[ob], tb, 4 * kGuideWidth + ob * 16, lane);
    }
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) store_transposed(h[3][kb], lds, at + (25 + kb) * kFragHalves, lane);
    // back through layers 4..1 (indices 3..0)
#pragma unroll
    for (int l = 3; l >= 0; --l) {
        const int ob_count = (l == 3) ? 2 : 4;
        float4_t d[4];
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) {
            d[ib] = zero4();
#pragma unroll
            for (int ob = 0; ob < 4; ++ob)
                if (ob < ob_count)
                    d[ib] = mfma16(Wb[(guide_bwd_base(l + 1) + ib * ob_count + ob) * 64 + lane], zh[ob], d[ib]);
        }
        // dZ_l = dH_l * (h_l > 0)
#pragma unroll
        for (int ib = 0; ib < 4; ++ib)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (!((float)h[l][ib][c] > 0.0f)) d[ib][c] = 0.0f;
#pragma unroll
        for (int ib = 0; ib < 4; ++ib) {
            zh[ib] = to_half4(d[ib]);
            store_transposed(zh[ib], lds, dz + (4 * l + ib) * kFragHalves, lane);
            add_bias_grad(d[ib], tb, l * kGuideWidth + ib * 16, lane);
        }
        if (l > 0) {
#pragma unroll
            for (int kb = 0; kb < 4; ++kb)
                store_transposed(h[l - 1][kb], lds, at + (13 + 4 * (l - 1) + kb) * kFragHalves, lane);
        }
    }
    // layer-1 inputs as dW A fragments: lane holds feature kb*16 + lane%16 of
    // queries 4(lane/16)..+3, read straight from the feature tile
#pragma unroll
    for (int kb = 0; kb < 13; ++kb) {
        half4_t r;
#pragma unroll
        for (int c = 0; c < 4; ++c) r[c] = ft[(4 * g + c) * kFeatStride + kb * 16 + (lane & 15u)];
        reinterpret_cast<half4_t*>(at + kb * kFragHalves)[lane] = r;
    }
}

// dW_l[in][out] = sum over split y's tiles of A_{l-1}[q][in] dZ_l[q][out]: one
// 16x16 output tile per block, the batch split kDwSplit ways (tiles y, y +
// kDwSplit, ...), each split's sum stored to part[y].
__global__ __launch_bounds__(64) void k_guide_dw(const uint16_t* __restrict__ at_buf,
                                                 const uint16_t* __restrict__ dz_buf, uint32_t tiles,
                                                 float* __restrict__ part) {
    const uint32_t lane = threadIdx.x;
    const int id = (int)blockIdx.x;
    int l, ib, ob;
    if (id < 52) {
        l = 0, ib = id / 4, ob = id % 4;
    } else if (id < 100) {
        l = 1 + (id - 52) / 16;
        ib = ((id - 52) % 16) / 4;
        ob = (id - 52) % 4;
    } else {
        l = 4, ib = (id - 100) / 2, ob = (id - 100) % 2;
    }
    const int a_off = l == 0 ? ib : 13 + 4 * (l - 1) + ib;
    const int z_off = 4 * l + ob;
    float4_t acc = zero4();
    for (uint32_t t = blockIdx.y; t < tiles; t += gridDim.y) {
        const half4_t a = reinterpret_cast<const half4_t*>(at_buf + ((size_t)t * kAFrags + a_off) * kFragHalves)[lane];
        const half4_t z = reinterpret_cast<const half4_t*>(dz_buf + ((size_t)t * kZFrags + z_off) * kFragHalves)[lane];
        acc = mfma16(a, z, acc);
    }
    const int nout = guide_layer_out(l);
    const int out = ob * 16 + (int)(lane & 15u);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int in = ib * 16 + 4 * (int)(lane >> 4) + c;
        part[(size_t)blockIdx.y * kGuideParams + guide_w_off(l) + in * nout + out] = acc[c];
    }
}

// Bias gradients of split y: the sum of its tiles' partials, in tile order.
__global__ __launch_bounds__(64) void k_guide_bias_reduce(const float* __restrict__ bias_part, uint32_t tiles,
                                                          float* __restrict__ part) {
    const int b = (int)(blockIdx.x * 64 + threadIdx.x);
    if (b >= kGuideBiases) return;
    float s = 0.0f;
    for (uint32_t t = blockIdx.y; t < tiles; t += gridDim.y) s += bias_part[(size_t)t * kGuideBiases + b];
    part[(size_t)blockIdx.y * kGuideParams + bias_param(b)] = s;
}

// grad += the splits' partial gradients, in split order.
__global__ __launch_bounds__(256) void k_guide_grad_reduce(const float* __restrict__ part, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kGuideParams) return;
    float g = grad[p];
#pragma unroll
    for (int y = 0; y < kDwSplit; ++y) g += part[(size_t)y * kGuideParams + p];
    grad[p] = g;
}

__device__ __forceinline__ void store_fragments(uint16_t* fwd, uint16_t* bwd, uint16_t* fwd32, int p, float w) {
    int l, in, out;
    if (!guide_weight_coords(p, &l, &in, &out)) return;
    const _Float16 h = (_Float16)w;
    uint16_t bits;
    __builtin_memcpy(&bits, &h, 2);
    fwd[guide_fwd_index(l, in, out)] = bits;
    const int i32 = guide_fwd32_index(l, in, out);
    if (i32 >= 0) fwd32[i32] = bits;
    if (l >= 1) bwd[guide_bwd_index(l, in, out)] = bits;
}

__global__ __launch_bounds__(256) void k_guide_pack(const float* __restrict__ params, uint16_t* __restrict__ fwd,
                                                    uint16_t* __restrict__ bwd, uint16_t* __restrict__ fwd32) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kGuideParams) store_fragments(fwd, bwd, fwd32, p, params[p]);
}

// Adam on the mean gradient of `count` records; clears the gradient.  Skipped
// (gradient still cleared) when no record was available.
__global__ __launch_bounds__(256) void k_guide_adam(float* __restrict__ params, float* __restrict__ m,
                                                    float* __restrict__ v, float* __restrict__ grad,
                                                    uint16_t* __restrict__ fwd, uint16_t* __restrict__ bwd,
                                                    uint16_t* __restrict__ fwd32,
                                                    const uint32_t* __restrict__ n_ptr, uint32_t n_host,
                                                    float inv_batch, float lr, float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kGuideParams) return;
    const float gr = grad[p] * inv_batch;
    grad[p] = 0.0f;
    const uint32_t n = n_ptr ? *n_ptr : n_host;
    if (n == 0) return;
    const float mm = 0.9f * m[p] + 0.1f * gr;
    const float vv = 0.999f * v[p] + 0.001f * (gr * gr);
    m[p] = mm;
    v[p] = vv;
    const float w = params[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + 1.0e-8f);
    params[p] = w;
    store_fragments(fwd, bwd, fwd32, p, w);
}

__global__ __launch_bounds__(256) void k_guide_fill(float* __restrict__ a, uint32_t n, float value) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = value;
}

// Per-block sums of the first n record values (a fixed tree per block).
__global__ __launch_bounds__(256) void k_guide_value_sum(const float* __restrict__ recs, uint32_t n_max,
                                                         const uint32_t* __restrict__ n_ptr,
                                                         float* __restrict__ vpart) {
    __shared__ float s[256];
    uint32_t n = n_max;
    if (n_ptr) {
        const uint32_t m = *n_ptr;
        n = m < n_max ? m : n_max;
    }
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    s[threadIdx.x] = i < n ? fmaxf(recs[(size_t)i * kRecordFloats + 11], 0.0f) : 0.0f;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) vpart[blockIdx.x] = s[0];
}

// stats[0] += the sum of the block sums (one block, a fixed order).
__global__ __launch_bounds__(256) void k_guide_value_total(const float* __restrict__ vpart, uint32_t nblocks,
                                                           float* __restrict__ stats) {
    __shared__ float s[256];
    float run = 0.0f;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256u) run += vpart[b];
    s[threadIdx.x] = run;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) stats[0] += s[0];
}

__global__ __launch_bounds__(256) void k_guide_scale_copy(const float* __restrict__ grad, float inv,
                                                          float* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kGuideParams) out[p] = grad[p] * inv;
}

__global__ __launch_bounds__(256) void k_mixture_sample(const float* __restrict__ mix, const float* __restrict__ u,
                                                        uint32_t n, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* m = mix + (size_t)i * kMixFloats;
    const vec3 w = mixture_sample(m, u[3 * (size_t)i], u[3 * (size_t)i + 1], u[3 * (size_t)i + 2]);
    out[4 * (size_t)i + 0] = w.x;
    out[4 * (size_t)i + 1] = w.y;
    out[4 * (size_t)i + 2] = w.z;
    out[4 * (size_t)i + 3] = mixture_pdf(m, w);
}

__global__ __launch_bounds__(256) void k_mixture_pdf(const float* __restrict__ mix, const float* __restrict__ dirs,
                                                     uint32_t n, float* __restrict__ pdf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pdf[i] = mixture_pdf(mix + (size_t)i * kMixFloats, mk(dirs[3 * (size_t)i], dirs[3 * (size_t)i + 1],
                                                             dirs[3 * (size_t)i + 2]));
}

inline uint32_t blocks256(size_t n) { return (uint32_t)((n + 255) / 256); }

int cus_of_device() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            cus = v;
        else
            return 64;
    }
    return cus;
}

hipError_t reserve_work(gr_guide* g, uint32_t batch) {
    const uint32_t tiles = (batch + 15) / 16;
    if (tiles <= g->work_tiles) return hipSuccess;
    (void)hipFree(g->work);
    g->work = nullptr;
    g->work_tiles = 0;
    const size_t a = (size_t)tiles * kAFrags * kFragHalves * 2, z = (size_t)tiles * kZFrags * kFragHalves * 2;
    const size_t bp = (size_t)tiles * kGuideBiases * 4, pt = (size_t)kDwSplit * kGuideParams * 4;
    char* p = nullptr;
    hipError_t e = hipMalloc(&p, a + z + bp + pt);
    if (e != hipSuccess) return e;
    g->work = p;
    g->at = reinterpret_cast<uint16_t*>(p);
    g->dz = reinterpret_cast<uint16_t*>(p + a);
    g->bias_part = reinterpret_cast<float*>(p + a + z);
    g->part = reinterpret_cast<float*>(p + a + z + bp);
    g->work_tiles = tiles;
    return hipSuccess;
}

// Forward/backward of `batch` records plus the weight-gradient reduction into g->grad.
hipError_t accumulate_gradient(gr_guide* g, const float* recs, uint32_t n_max, const uint32_t* n_ptr,
                               uint32_t batch, uint32_t draw_key, bool sequential, hipStream_t s) {
    hipError_t e = reserve_work(g, batch);
    if (e != hipSuccess) return e;
    const uint32_t tiles = (batch + 15) / 16;
    const uint32_t blocks = (tiles + kTrainThreads / 64 - 1) / (kTrainThreads / 64);
    hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(kTrainThreads), 0, s, g->fwd, g->bwd, g->params,
                       g->bias_part, recs, n_max, n_ptr, g->stats, batch, draw_key, sequential ? 1u : 0u, g->at,
                       g->dz);
    hipLaunchKernelGGL(k_guide_dw, dim3(kDwTiles, kDwSplit), dim3(64), 0, s, g->at, g->dz, tiles, g->part);
    hipLaunchKernelGGL(k_guide_bias_reduce, dim3((kGuideBiases + 63) / 64, kDwSplit), dim3(64), 0, s,
                       g->bias_part, tiles, g->part);
    hipLaunchKernelGGL(k_guide_grad_reduce, dim3(blocks256(kGuideParams)), dim3(256), 0, s, g->part, g->grad);
    return hipGetLastError();
}

// SplitMix64, for the weight initialisation.
uint64_t splitmix(uint64_t* s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// He-uniform weights (bound sqrt(6 / fan_in)), the output layer scaled by 0.1 so
// the first mixture is close to uniform; zero biases and zero padding rows.
std::vector<float> initial_params(uint64_t seed) {
    std::vector<float> p(kGuideParams, 0.0f);
    uint64_t st = seed ^ 0x6a09e667f3bcc909ull;
    for (int l = 0; l < 5; ++l) {
        const int nin = guide_layer_in(l), nout = guide_layer_out(l);
        const int fan_in = l == 0 ? kGuideEncoded + kGuideRaw : nin;
        const float bound = (l == 4 ? 0.1f : 1.0f) * sqrtf(6.0f / (float)fan_in);
        for (int i = 0; i < nin; ++i)
            for (int o = 0; o < nout; ++o) {
                const float u = (float)(splitmix(&st) >> 40) * (1.0f / 16777216.0f);
                p[guide_w_off(l) + i * nout + o] = (l == 0 && i >= fan_in) ? 0.0f : (2.0f * u - 1.0f) * bound;
            }
    }
    return p;
}

int set_error(gr_ctx* ctx, int code, const char* what) {
    if (ctx) ctx->err = what;
    return code;
}

int hip_error(gr_ctx* ctx, hipError_t e, const char* what) {
    if (ctx) ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return GR_ERR_HIP;
}

hipError_t upload_params(gr_guide* g, const float* host) {
    hipStream_t s = g->ctx->stream;
    const size_t bytes = (size_t)kGuideParams * 4;
    hipError_t e = hipMemcpyAsync(g->params, host, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(g->m, 0, bytes, s);
    if (e == hipSuccess) e = hipMemsetAsync(g->v, 0, bytes, s);
    if (e == hipSuccess) e = hipMemsetAsync(g->grad, 0, bytes, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_guide_pack, dim3(blocks256(kGuideParams)), dim3(256), 0, s, g->params, g->fwd, g->bwd,
                           g->fwd32);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    g->steps = 0;
    return e;
}

}  // namespace

// ---- internal interface ------------------------------------------------------
hipError_t guide_infer(gr_guide* g, const float* raw, uint32_t n_max, const uint32_t* n_ptr, float* mix,
                       hipStream_t s) {
    if (n_max == 0) return hipSuccess;
    const uint32_t spans = (n_max + 16 * kInferTiles - 1) / (16 * kInferTiles);
    uint32_t blocks = (spans + kInferThreads / 64 - 1) / (kInferThreads / 64);
    // two blocks per CU: each holds its own copy of the weights (57 KiB of LDS)
    const uint32_t cap = 2u * (uint32_t)cus_of_device();
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_guide_infer, dim3(blocks), dim3(kInferThreads), 0, s, g->fwd32, g->params, raw, n_max,
                       n_ptr, mix);
    return hipGetLastError();
}

hipError_t guide_zero_stats(gr_guide* g, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_fill, dim3(1), dim3(256), 0, s, g->stats, 4u, 0.0f);
    return hipGetLastError();
}

hipError_t guide_record_stats(gr_guide* g, const float* recs, uint32_t n_max, const uint32_t* n_ptr, hipStream_t s) {
    if (n_max == 0) return hipSuccess;
    const uint32_t nb = blocks256(n_max);
    if (nb > g->vpart_cap) {
        (void)hipFree(g->vpart);
        g->vpart = nullptr;
        g->vpart_cap = 0;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&g->vpart), (size_t)nb * sizeof(float));
        if (e != hipSuccess) return e;
        g->vpart_cap = nb;
    }
    hipLaunchKernelGGL(k_guide_value_sum, dim3(nb), dim3(256), 0, s, recs, n_max, n_ptr, g->vpart);
    hipLaunchKernelGGL(k_guide_value_total, dim3(1), dim3(256), 0, s, g->vpart, nb, g->stats);
    return hipGetLastError();
}

hipError_t guide_adam_steps(gr_guide* g, const float* recs, uint32_t n_max, const uint32_t* n_ptr, uint32_t steps,
                            hipStream_t s) {
    const uint32_t batch = g->prm.batch;
    for (uint32_t k = 0; k < steps; ++k) {
        const uint32_t key = (uint32_t)(g->prm.seed * 0x9e3779b1u) ^ (uint32_t)(g->steps * 0x85ebca6bu) ^ 0x5bd1e995u;
        hipError_t e = accumulate_gradient(g, recs, n_max, n_ptr, batch, key, false, s);
        if (e != hipSuccess) return e;
        g->steps += 1;
        const double t = (double)g->steps;
        const float bc1 = (float)(1.0 - pow(0.9, t)), bc2 = (float)(1.0 - pow(0.999, t));
        hipLaunchKernelGGL(k_guide_adam, dim3(blocks256(kGuideParams)), dim3(256), 0, s, g->params, g->m, g->v,
                           g->grad, g->fwd, g->bwd, g->fwd32, n_ptr, n_max, 1.0f / (float)batch,
                           g->prm.learning_rate, bc1, bc2);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t guide_reserve_records(gr_guide* g, size_t n) {
    if (n <= g->recs_cap) return hipSuccess;
    (void)hipFree(g->recs_blob);
    g->recs_blob = nullptr;
    g->recs_cap = 0;
    const size_t nb = (n + 255) / 256, rec_bytes = (n * kRecordFloats * sizeof(float) + 255) & ~(size_t)255;
    hipError_t e = hipMalloc(&g->recs_blob, rec_bytes + 2 * nb * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    g->recs = reinterpret_cast<float*>(g->recs_blob);
    g->rec_bcount = reinterpret_cast<uint32_t*>(static_cast<char*>(g->recs_blob) + rec_bytes);
    g->rec_boff = g->rec_bcount + nb;
    g->recs_cap = n;
    return hipSuccess;
}

}  // namespace gr

// ---- C ABI ------------------------------------------------------------------
extern "C" {

int gr_guide_create(gr_ctx* ctx, const gr_guide_params* prm, gr_guide** out) {
    if (!ctx || !prm || !out) return GR_ERR_INVALID;
    *out = nullptr;
    if (ctx->device == GR_HOST_ONLY) return gr::set_error(ctx, GR_ERR_STATE, "host-only context");
    if (prm->batch == 0 || prm->batch % 16 != 0 || prm->batch > (1u << 22))
        return gr::set_error(ctx, GR_ERR_INVALID, "guide batch must be a positive multiple of 16 (<= 4M)");
    if (!(prm->alpha >= 0.0f && prm->alpha < 1.0f)) return gr::set_error(ctx, GR_ERR_INVALID, "alpha must be in [0, 1)");
    if (prm->lobes != GR_GUIDE_LOBES_CONSISTENT && prm->lobes != GR_GUIDE_LOBES_ALL &&
        prm->lobes != GR_GUIDE_LOBES_DIFFUSE)
        return gr::set_error(ctx, GR_ERR_INVALID, "unknown guided lobe set");
    if (prm->lobe_select != GR_GUIDE_SELECT_UNIFORM && prm->lobe_select != GR_GUIDE_SELECT_ALBEDO &&
        prm->lobe_select != GR_GUIDE_SELECT_LEARNED)
        return gr::set_error(ctx, GR_ERR_INVALID, "unknown lobe choice");
    if (prm->lobe_select != GR_GUIDE_SELECT_UNIFORM && !(prm->select_mix > 0.0f && prm->select_mix <= 1.0f))
        return gr::set_error(ctx, GR_ERR_INVALID, "select_mix must be in (0, 1]");
    if (!(prm->rr_threshold >= 0.0f && prm->rr_threshold < 1.0e30f))
        return gr::set_error(ctx, GR_ERR_INVALID, "rr_threshold must be finite and >= 0");
    if (!(prm->learning_rate > 0.0f)) return gr::set_error(ctx, GR_ERR_INVALID, "learning rate must be positive");
    gr_guide* g = new (std::nothrow) gr_guide();
    if (!g) return gr::set_error(ctx, GR_ERR_NOMEM, "out of host memory");
    g->ctx = ctx;
    g->prm = *prm;
    const size_t P = gr::kGuideParams;
    const size_t f32 = ((P * 4 + 255) / 256) * 256;
    const size_t fwd = (size_t)gr::kFwdTiles * gr::kFragHalves * 2, bwd = (size_t)gr::kBwdTiles * gr::kFragHalves * 2;
    const size_t fwd32 = (size_t)gr::kFwd32Tiles * gr::kFrag32Halves * 2;
    char* p = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMalloc(&p, 4 * f32 + fwd + bwd + 256 + fwd32);
    if (e != hipSuccess) {
        delete g;
        return gr::hip_error(ctx, e, "guide allocation");
    }
    g->blob = p;
    g->params = reinterpret_cast<float*>(p);
    g->m = reinterpret_cast<float*>(p + f32);
    g->v = reinterpret_cast<float*>(p + 2 * f32);
    g->grad = reinterpret_cast<float*>(p + 3 * f32);
    g->fwd = reinterpret_cast<uint16_t*>(p + 4 * f32);
    g->bwd = reinterpret_cast<uint16_t*>(p + 4 * f32 + fwd);
    g->stats = reinterpret_cast<float*>(p + 4 * f32 + fwd + bwd);
    g->count = reinterpret_cast<uint32_t*>(p + 4 * f32 + fwd + bwd + 128);
    g->fwd32 = reinterpret_cast<uint16_t*>(p + 4 * f32 + fwd + bwd + 256);
    // zero: the fragments' padding (layer 1's inputs 208..223 in fwd32) is never written
    e = hipMemsetAsync(p + 4 * f32, 0, fwd + bwd + 256 + fwd32, ctx->stream);
    const std::vector<float> init = gr::initial_params(prm->seed);
    if (e == hipSuccess) e = gr::upload_params(g, init.data());
    if (e != hipSuccess) {
        (void)hipFree(g->blob);
        delete g;
        return gr::hip_error(ctx, e, "guide initialisation");
    }
    gr::ctx_retain(ctx);
    *out = g;
    return GR_OK;
}

void gr_guide_destroy(gr_guide* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    (void)hipFree(g->blob);
    (void)hipFree(g->work);
    (void)hipFree(g->recs_blob);
    (void)hipFree(g->vpart);
    (void)hipFree(g->lobe_blob);
    gr_ctx* ctx = g->ctx;
    delete g;
    gr::ctx_release(ctx);
}

int gr_guide_get_params(gr_guide* g, float* host, uint32_t n) {
    if (!g || !host) return GR_ERR_INVALID;
    if (n != (uint32_t)gr::kGuideParams) return gr::set_error(g->ctx, GR_ERR_INVALID, "parameter count mismatch");
    hipError_t e = hipMemcpyAsync(host, g->params, (size_t)n * 4, hipMemcpyDeviceToHost, g->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
    return e == hipSuccess ? GR_OK : gr::hip_error(g->ctx, e, "guide parameters");
}

int gr_guide_set_params(gr_guide* g, const float* host, uint32_t n) {
    if (!g || !host) return GR_ERR_INVALID;
    if (n != (uint32_t)gr::kGuideParams) return gr::set_error(g->ctx, GR_ERR_INVALID, "parameter count mismatch");
    const hipError_t e = gr::upload_params(g, host);
    return e == hipSuccess ? GR_OK : gr::hip_error(g->ctx, e, "guide parameters");
}

int gr_guide_steps(gr_guide* g, uint64_t* steps) {
    if (!g || !steps) return GR_ERR_INVALID;
    *steps = g->steps;
    return GR_OK;
}

int gr_guide_infer(gr_guide* g, const float* d_raw, uint32_t n, float* d_mix) {
    if (!g) return GR_ERR_INVALID;
    if (n && (!d_raw || !d_mix)) return gr::set_error(g->ctx, GR_ERR_INVALID, "null buffer");
    if ((reinterpret_cast<uintptr_t>(d_raw) | reinterpret_cast<uintptr_t>(d_mix)) & 15u)
        return gr::set_error(g->ctx, GR_ERR_INVALID, "raw inputs and mixtures must be 16-byte aligned");
    const hipError_t e = gr::guide_infer(g, d_raw, n, nullptr, d_mix, g->ctx->stream);
    return e == hipSuccess ? GR_OK : gr::hip_error(g->ctx, e, "guide inference");
}

int gr_guide_gradient(gr_guide* g, const float* d_records, uint32_t n, float* d_grad) {
    if (!g) return GR_ERR_INVALID;
    if (n == 0 || !d_records || !d_grad) return gr::set_error(g->ctx, GR_ERR_INVALID, "need records and an output");
    hipStream_t s = g->ctx->stream;
    const uint32_t batch = (n + 15) / 16 * 16;
    hipError_t e = gr::guide_zero_stats(g, s);
    if (e == hipSuccess) e = gr::guide_record_stats(g, d_records, n, nullptr, s);
    if (e == hipSuccess) e = gr::accumulate_gradient(g, d_records, n, nullptr, batch, 0u, true, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gr::k_guide_scale_copy, dim3(gr::blocks256(gr::kGuideParams)), dim3(256), 0, s, g->grad,
                           1.0f / (float)n, d_grad);
        hipLaunchKernelGGL(gr::k_guide_fill, dim3(gr::blocks256(gr::kGuideParams)), dim3(256), 0, s, g->grad,
                           (uint32_t)gr::kGuideParams, 0.0f);
        e = hipGetLastError();
    }
    return e == hipSuccess ? GR_OK : gr::hip_error(g->ctx, e, "guide gradient");
}

int gr_guide_train(gr_guide* g, const float* d_records, uint32_t n, uint32_t steps) {
    if (!g) return GR_ERR_INVALID;
    if (n == 0 || !d_records) return gr::set_error(g->ctx, GR_ERR_INVALID, "need records");
    hipStream_t s = g->ctx->stream;
    hipError_t e = gr::guide_zero_stats(g, s);
    if (e == hipSuccess) e = gr::guide_record_stats(g, d_records, n, nullptr, s);
    if (e == hipSuccess) e = gr::guide_adam_steps(g, d_records, n, nullptr, steps, s);
    return e == hipSuccess ? GR_OK : gr::hip_error(g->ctx, e, "guide training");
}

int gr_guide_last_records(gr_guide* g, float* d_out, uint32_t cap, uint32_t* n_out) {
    if (!g || !n_out || (cap && !d_out)) return GR_ERR_INVALID;
    hipStream_t s = g->ctx->stream;
    uint32_t count = 0;
    hipError_t e = g->recs ? hipMemcpyAsync(&count, g->count, sizeof(count), hipMemcpyDeviceToHost, s) : hipSuccess;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return gr::hip_error(g->ctx, e, "record count");
    uint32_t n = g->recs ? count : 0u;
    n = std::min<uint32_t>(n, (uint32_t)std::min<size_t>(g->recs_cap, 0xffffffffu));
    n = std::min(n, cap);
    if (n) e = hipMemcpyAsync(d_out, g->recs, (size_t)n * gr::kRecordFloats * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return gr::hip_error(g->ctx, e, "record copy");
    *n_out = n;
    return GR_OK;
}

int gr_guide_mixture_sample(gr_ctx* ctx, const float* d_mix, const float* d_u, uint32_t n, float* d_out) {
    if (!ctx) return GR_ERR_INVALID;
    if (n == 0) return GR_OK;
    if (!d_mix || !d_u || !d_out) return gr::set_error(ctx, GR_ERR_INVALID, "null buffer");
    hipLaunchKernelGGL(gr::k_mixture_sample, dim3(gr::blocks256(n)), dim3(256), 0, ctx->stream, d_mix, d_u, n, d_out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? GR_OK : gr::hip_error(ctx, e, "mixture sampling");
}

int gr_guide_mixture_pdf(gr_ctx* ctx, const float* d_mix, const float* d_dirs, uint32_t n, float* d_pdf) {
    if (!ctx) return GR_ERR_INVALID;
    if (n == 0) return GR_OK;
    if (!d_mix || !d_dirs || !d_pdf) return gr::set_error(ctx, GR_ERR_INVALID, "null buffer");
    hipLaunchKernelGGL(gr::k_mixture_pdf, dim3(gr::blocks256(n)), dim3(256), 0, ctx->stream, d_mix, d_dirs, n, d_pdf);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? GR_OK : gr::hip_error(ctx, e, "mixture pdf");
}

}  // extern "C"
