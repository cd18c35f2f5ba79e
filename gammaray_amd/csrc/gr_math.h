// gr_math.h -- f32 vector math shared by the host builder and the gfx950
// kernels.  Operation order follows src/core/vector.rs and src/core/math.rs of
// the reference; every translation unit is compiled with -ffp-contract=off so
// a*b+c is never fused (Rust never fuses), which keeps traversal bit-exact.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstdint>

#define GR_HD __host__ __device__ __forceinline__

namespace gr {

constexpr float kEps = FLT_EPSILON;          // std::f32::EPSILON
constexpr float kRayPush = 1.0e-3f;          // RAY_PUSH_DIST, src/core/ray.rs:9
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr uint32_t kNone = 0xffffffffu;

struct vec3 {
    float x, y, z;
};

GR_HD vec3 mk(float x, float y, float z) { return vec3{x, y, z}; }
GR_HD vec3 operator+(vec3 a, vec3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
GR_HD vec3 operator-(vec3 a, vec3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
GR_HD vec3 operator-(vec3 a) { return mk(-a.x, -a.y, -a.z); }
GR_HD vec3 operator*(vec3 a, float k) { return mk(a.x * k, a.y * k, a.z * k); }
GR_HD vec3 operator/(vec3 a, float k) { return mk(a.x / k, a.y / k, a.z / k); }
GR_HD vec3 mul(vec3 a, vec3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
GR_HD float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GR_HD vec3 cross(vec3 a, vec3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
GR_HD float length(vec3 a) { return sqrtf(dot(a, a)); }
GR_HD vec3 normalized(vec3 a) {
    float l = length(a);
    return mk(a.x / l, a.y / l, a.z / l);
}
GR_HD float comp(vec3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
GR_HD bool nearly_zero(float x) { return fabsf(x) < kEps; }   // math.rs:41-43
GR_HD bool positive(float x) { return x > kEps; }             // math.rs:48-50

// 1 + 2 * gamma(3) (src/core/math.rs:152-154, src/core/bbox.rs:88) -- the
// slab test's t_max padding, evaluated in f32 exactly as the reference does.
inline float slab_pad() {
    volatile float n = 3.0f;
    float g = (n * FLT_EPSILON) / (1.0f - (n * FLT_EPSILON));
    return 1.0f + 2.0f * g;
}

}  // namespace gr
