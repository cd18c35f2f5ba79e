// gr_guide.h -- device side of the learned path guide: input encoding,
// parameter layout, MFMA fragment layout of the fp16 weights, and the
// von Mises-Fisher mixture the network outputs.
//
// Network: 8 raw inputs (position in the scene box, outgoing direction and
// shading normal as (theta/pi, phi/2pi + 1/2), roughness), each encoded as
// sin/cos(2^k pi x) for k = 0..11 (192 features) plus the 8 raw values, padded
// to 208; then 208 -> 64 -> 64 -> 64 -> 64 (ReLU) -> 32 outputs = 8 lobes x
// (weight logit, log kappa, theta, phi).  27,936 parameters.
#pragma once
#pragma clang fp contract(off)

#include "gr_math.h"

namespace gr {

constexpr int kGuideRaw = 8;
constexpr int kGuideFreqs = 12;
constexpr int kGuideEncoded = kGuideRaw * 2 * kGuideFreqs;  // 192
constexpr int kGuideIn = 208;                               // 192 + 8 raw + 8 zero
constexpr int kGuideWidth = 64;
constexpr int kGuideOut = 32;
constexpr int kGuideLobes = 8;
constexpr int kMixFloats = 5 * kGuideLobes;  // per query: lobe i -> (w, kappa, mu.x, mu.y, mu.z)
constexpr int kRecordFloats = 12;            // training record: raw inputs(8), direction(3), value

// fp32 parameter vector: W_l row-major [in][out], then b_l.
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kGuideIn * kGuideWidth;
constexpr int kW2 = kB1 + kGuideWidth;
constexpr int kB2 = kW2 + kGuideWidth * kGuideWidth;
constexpr int kW3 = kB2 + kGuideWidth;
constexpr int kB3 = kW3 + kGuideWidth * kGuideWidth;
constexpr int kW4 = kB3 + kGuideWidth;
constexpr int kB4 = kW4 + kGuideWidth * kGuideWidth;
constexpr int kW5 = kB4 + kGuideWidth;
constexpr int kB5 = kW5 + kGuideWidth * kGuideOut;
constexpr int kGuideParams = kB5 + kGuideOut;  // 27,936

// MFMA fragments (v_mfma_f32_16x16x16_f16): a 16x16 operand tile is 64 lanes x
// 4 halves.  A operand: lane L holds row L%16, k = 4(L/16)..+3.  B operand:
// lane L holds column L%16, k = 4(L/16)..+3.  C/D: lane L holds column L%16,
// rows 4(L/16)..+3.  Activations are kept transposed (rows = features,
// columns = 16 queries), so a layer's C tiles are the next layer's B tiles.
constexpr int kFragHalves = 256;
// Forward fragments: A = W^T tiles (rows = out, k = in), per layer [out tile][in tile];
// layer bases 0, 52, 68, 84, 100.
constexpr int kFwdTiles = 108;
__host__ __device__ constexpr int guide_fwd_base(int l) { return l == 0 ? 0 : 52 + 16 * (l - 1); }
// Backward fragments: A = W tiles (rows = in, k = out), layers 2..5, per layer
// [in tile][out tile]; layer bases 0, 16, 32, 48.
constexpr int kBwdTiles = 56;
__host__ __device__ constexpr int guide_bwd_base(int l) { return 16 * (l - 1); }

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Inference fragments for v_mfma_f32_16x16x32_f16 (A: lane L holds row L%16,
// k = 8(L/16)..+7; B: column L%16, same k; C as for 16x16x16).  Layer 1's k
// runs over 7 blocks of 32 in the order the encoder produces features in
// registers: lane group G (lane / 16) owns input dimensions 2G and 2G+1, and
// its 8 values of block kb are features 48G + 8kb .. +7 for kb < 6 (the
// sin/cos pairs of one dimension at 4 consecutive frequencies) and, for kb =
// 6, raw inputs 2G, 2G+1 and six zeros.  A hidden layer's 32-wide k block kb
// is the pair of C tiles 2kb, 2kb+1 as each lane holds them -- k = 8G + i is
// feature (2kb + i/4)*16 + 4G + i%4 -- so activations move straight from one
// layer's C tiles into the next layer's B operand.
constexpr int kFrag32Halves = 512;
constexpr int kFwd32Tiles = 56;  // 4x7, 4x2, 4x2, 4x2, 2x2
__host__ __device__ constexpr int guide_fwd32_base(int l) { return l == 0 ? 0 : 28 + 8 * (l - 1); }
// Fragment position of weight (l, in, out); -1 for layer 1's zero padding rows.
__host__ __device__ inline int guide_fwd32_index(int l, int in, int out) {
    const int ot = out / 16;
    int kb, g, i;
    if (l == 0) {
        if (in < kGuideEncoded) {
            g = in / 48;
            kb = (in % 48) / 8;
            i = in % 8;
        } else if (in < kGuideEncoded + kGuideRaw) {
            g = (in - kGuideEncoded) / 2;
            kb = 6;
            i = (in - kGuideEncoded) % 2;
        } else {
            return -1;
        }
    } else {
        const int b16 = in / 16, r = in % 16;
        kb = b16 / 2;
        g = r / 4;
        i = (b16 % 2) * 4 + r % 4;
    }
    const int tile = guide_fwd32_base(l) + ot * (l == 0 ? 7 : 2) + kb;
    return tile * kFrag32Halves + (g * 16 + out % 16) * 8 + i;
}

__host__ __device__ inline int guide_layer_in(int l) { return l == 0 ? kGuideIn : kGuideWidth; }
__host__ __device__ inline int guide_layer_out(int l) { return l == 4 ? kGuideOut : kGuideWidth; }
__host__ __device__ inline int guide_w_off(int l) {
    return l == 0 ? kW1 : (l == 1 ? kW2 : (l == 2 ? kW3 : (l == 3 ? kW4 : kW5)));
}
__host__ __device__ inline int guide_b_off(int l) {
    return l == 0 ? kB1 : (l == 1 ? kB2 : (l == 2 ? kB3 : (l == 3 ? kB4 : kB5)));
}

// Parameter index -> (layer, in, out) for weights; false for biases.
__host__ __device__ inline bool guide_weight_coords(int p, int* layer, int* in, int* out) {
    for (int l = 0; l < 5; ++l) {
        const int w = guide_w_off(l), n = guide_layer_out(l);
        if (p >= w && p < guide_b_off(l)) {
            *layer = l;
            *in = (p - w) / n;
            *out = (p - w) % n;
            return true;
        }
    }
    return false;
}

// Half index of W_l[in][out] in the forward / backward fragment arrays.
__host__ __device__ inline int guide_fwd_index(int l, int in, int out) {
    const int kb_count = guide_layer_in(l) / 16;
    const int tile = guide_fwd_base(l) + (out / 16) * kb_count + in / 16;
    const int lane = ((in % 16) / 4) * 16 + out % 16;
    return tile * kFragHalves + lane * 4 + in % 4;
}
__host__ __device__ inline int guide_bwd_index(int l, int in, int out) {
    const int ob_count = guide_layer_out(l) / 16;
    const int tile = guide_bwd_base(l) + (in / 16) * ob_count + out / 16;
    const int lane = ((out % 16) / 4) * 16 + in % 16;
    return tile * kFragHalves + lane * 4 + out % 4;
}

// sin/cos(2^k pi x) = sin/cos(2 pi r) with r = 2^(k-1) x - floor(2^(k-1) x):
// the scaling and the reduction are exact in float, and v_sin/v_cos take the
// angle in revolutions (absolute error near 1e-6, far below the fp16
// rounding the features get).
__device__ __forceinline__ void guide_sincos(float x, int k, float* s, float* c) {
    const float a = ldexpf(x, k - 1);
    const float r = a - floorf(a);
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}

// Input encoding: sin/cos(2^k pi x) for k = 0..11 of each raw input, then
// the raw inputs themselves, then zeros to 208.  Each run of four frequencies
// k = a..a+3 (a = 0, 4, 8) starts from the exact values at k = a and takes
// three double-angle steps, so the error stays within 8x of the direct values.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Features 2a..2a+7 of input x, interleaved (sin, cos) per frequency.
__device__ __forceinline__ half8_t guide_encode_run(float x, int a) {
    float s, c;
    guide_sincos(x, a, &s, &c);
    half8_t v;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[2 * k] = (_Float16)s;
        v[2 * k + 1] = (_Float16)c;
        const float s2 = 2.0f * s * c;
        const float c2 = (c - s) * (c + s);
        s = s2;
        c = c2;
    }
    return v;
}

// Layer 1's K=32 B operand for block kb of the lane's query, whose inputs
// 2G and 2G+1 (G = lane / 16) are x0 and x1.
__device__ __forceinline__ half8_t guide_encode_block(float x0, float x1, int kb) {
    if (kb < 6) return guide_encode_run(kb < 3 ? x0 : x1, 4 * (kb % 3));
    half8_t v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (_Float16)0.0f;
    v[0] = (_Float16)x0;
    v[1] = (_Float16)x1;
    return v;
}

// Encoded features of 16 queries, staged in a per-wave LDS tile of 16 x
// kFeatStride halves (training reads it as K=16 B operands and, transposed,
// as the layer-1 activations of the weight gradient).  The lane holding query
// q and group g (lane / 16) writes input dimensions 2g and 2g+1; group 0
// also writes the zero padding.
constexpr int kFeatStride = 216;  // halves per query row: 208 used; 16-byte aligned rows

// xa, xb: the query's raw inputs 0..3 and 4..7.
__device__ __forceinline__ void guide_encode_tile(float4 xa, float4 xb, _Float16* tile, uint32_t lane) {
    const uint32_t q = lane & 15u, g = lane >> 4;
    _Float16* row = tile + q * kFeatStride;
    const float x0 = g == 0 ? xa.x : (g == 1 ? xa.z : (g == 2 ? xb.x : xb.z));
    const float x1 = g == 0 ? xa.y : (g == 1 ? xa.w : (g == 2 ? xb.y : xb.w));
#pragma unroll
    for (int r = 0; r < 6; ++r)
        *reinterpret_cast<half8_t*>(row + 48 * g + 8 * r) = guide_encode_block(x0, x1, r);
    half2_t raw;
    raw[0] = (_Float16)x0;
    raw[1] = (_Float16)x1;
    reinterpret_cast<half2_t*>(row + kGuideEncoded)[g] = raw;
    if (g == 0) {
        half2_t z;
        z[0] = z[1] = (_Float16)0.0f;
#pragma unroll
        for (int k = (kGuideEncoded + kGuideRaw) / 2; k < kGuideIn / 2; ++k) reinterpret_cast<half2_t*>(row)[k] = z;
    }
}

// Raw inputs of a surface vertex.  lo/inv_ext: scene box.
__device__ __forceinline__ float guide_angle01(float y, float x) {
    return atan2f(y, x) * (0.5f * kInvPi) + 0.5f;
}
__device__ __forceinline__ void guide_raw_inputs(vec3 p, vec3 wo, vec3 n, float roughness, const float lo[3],
                                                 const float inv_ext[3], float x[8]) {
    x[0] = (p.x - lo[0]) * inv_ext[0];
    x[1] = (p.y - lo[1]) * inv_ext[1];
    x[2] = (p.z - lo[2]) * inv_ext[2];
    x[3] = acosf(fmaxf(-1.0f, fminf(1.0f, wo.z))) * kInvPi;
    x[4] = guide_angle01(wo.y, wo.x);
    x[5] = acosf(fmaxf(-1.0f, fminf(1.0f, n.z))) * kInvPi;
    x[6] = guide_angle01(n.y, n.x);
    x[7] = roughness;
}

// ---- output head ------------------------------------------------------------
constexpr float kLogKappaLo = -4.0f, kLogKappaHi = 7.0f;

__device__ __forceinline__ float guide_sigmoid(float v) { return 1.0f / (1.0f + expf(-v)); }

__device__ __forceinline__ vec3 guide_mu(float theta, float phi) {
    const float st = sinf(theta);
    return mk(st * cosf(phi), st * sinf(phi), cosf(theta));
}

// vMF density kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1)).
__device__ __forceinline__ float vmf_pdf(float kappa, vec3 mu, vec3 w) {
    const float t = -expm1f(-2.0f * kappa);
    return kappa / (2.0f * kPi * t) * expf(kappa * (dot(mu, w) - 1.0f));
}

__device__ __forceinline__ float mixture_pdf(const float* mix, vec3 w) {
    float p = 0.0f;
    for (int i = 0; i < kGuideLobes; ++i) {
        const float* l = mix + 5 * i;
        p += l[0] * vmf_pdf(l[1], mk(l[2], l[3], l[4]), w);
    }
    return p;
}

// Lobe by u0 (inverse CDF of the weights), then the vMF by inversion of its
// cosine distribution with a uniform azimuth.
__device__ __forceinline__ vec3 mixture_sample(const float* mix, float u0, float u1, float u2) {
    int sel = kGuideLobes - 1;
    float acc = 0.0f;
    for (int i = 0; i < kGuideLobes; ++i) {
        acc += mix[5 * i];
        if (u0 < acc) {
            sel = i;
            break;
        }
    }
    const float* l = mix + 5 * sel;
    const float kappa = l[1];
    const vec3 mu = mk(l[2], l[3], l[4]);
    const float e = expf(-2.0f * kappa);
    float c = 1.0f + logf(e + u1 * (1.0f - e)) / kappa;
    c = fmaxf(-1.0f, fminf(1.0f, c));
    const float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    const float ph = 2.0f * kPi * u2;
    vec3 t, b;
    // orthonormal basis around mu (Duff et al. 2017)
    const float sg = mu.z >= 0.0f ? 1.0f : -1.0f;
    const float aa = -1.0f / (sg + mu.z);
    const float bb = mu.x * mu.y * aa;
    t = mk(1.0f + sg * mu.x * mu.x * aa, sg * bb, -sg * mu.x);
    b = mk(bb, sg + mu.y * mu.y * aa, -mu.y);
    return mu * c + (t * cosf(ph) + b * sinf(ph)) * s;
}

}  // namespace gr
