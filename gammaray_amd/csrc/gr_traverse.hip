// gr_traverse.hip -- closest-hit and shadow-ray traversal kernels for gfx950.
//
// Semantics are those of Bvh::intersect / Bvh::visibility
// (src/geom/bvh.rs:406-477) so that hits match the reference bit for bit:
//   * same binary tree, same near/far rule (dir_is_neg[axis], bvh.rs:437-445),
//   * the reference slab test (src/core/bbox.rs:78-113) with its t_max padding
//     and its z slab that never raises t_min,
//   * Moller-Trumbore against vertex c and the quadratic sphere test with the
//     reference's f32 operation order (src/geom/mesh.rs:270-304,
//     src/geom/sphere.rs:70-101),
//   * acceptance dist != 0 && dist < closest (first visited wins ties).
// Each 64-byte node carries both child boxes: a parent fetch tests both
// children, the near child is taken at once and the far child is pushed with
// its slab entry distance, which is re-compared with the then-current closest
// distance when popped -- the decision the reference makes when it pops and
// tests that child.  Box tests and primitive tests are therefore the same set,
// in the same order, as the reference's.
#pragma clang fp contract(off)

#include "gr_internal.h"

namespace gr {
namespace {

constexpr int kStackSize = 64;  // MAX_NODES_TO_VISIT, src/geom/bvh.rs:140

enum TraceMode { kClosest = 0, kOcclusionExact = 1, kOcclusionAnyHit = 2 };

struct RayFrame {
    vec3 o, d, inv;
    bool nx, ny, nz;
};

__device__ __forceinline__ RayFrame make_frame(vec3 o, vec3 d) {
    RayFrame r;
    r.o = o;
    r.d = d;
    // Ray::compute_intersection_data, src/core/ray.rs:31-42.
    r.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.nx = d.x < 0.0f;
    r.ny = d.y < 0.0f;
    r.nz = d.z < 0.0f;
    return r;
}

// The part of BBox::intersect that does not depend on the current closest
// distance; *tmin receives the entry distance the final `t_min < max_dist`
// compares.
__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                     const RayFrame& r, float pad, float* tmin) {
    float t_min = ((r.nx ? hx : lx) - r.o.x) * r.inv.x;
    float t_max = ((r.nx ? lx : hx) - r.o.x) * r.inv.x;
    float ty_min = ((r.ny ? hy : ly) - r.o.y) * r.inv.y;
    float ty_max = ((r.ny ? ly : hy) - r.o.y) * r.inv.y;
    t_max *= pad;
    ty_max *= pad;
    if (t_min > ty_max || ty_min > t_max) return false;
    if (ty_min > t_min) t_min = ty_min;
    if (ty_max < t_max) t_max = ty_max;
    float tz_min = ((r.nz ? hz : lz) - r.o.z) * r.inv.z;
    float tz_max = ((r.nz ? lz : hz) - r.o.z) * r.inv.z;
    tz_max *= pad;
    if (t_min > tz_max || tz_min > t_max) return false;
    if (tz_max < t_max) t_max = tz_max;
    *tmin = t_min;
    return t_max > 0.0f;
}

// Mesh::intersect_world distance (src/geom/mesh.rs:278-304) from the stored
// c, edge1 = a - c, edge2 = b - c.  Returns 0 on a miss.
__device__ __forceinline__ float hit_triangle(const GpuComp& g, const RayFrame& r) {
    const vec3 c = mk(g.q0.x, g.q0.y, g.q0.z);
    const vec3 e1 = mk(g.q0.w, g.q1.x, g.q1.y);
    const vec3 e2 = mk(g.q1.z, g.q1.w, g.q2.x);
    const vec3 p = cross(r.d, e2);
    const float det = dot(e1, p);
    if (nearly_zero(det)) return 0.0f;
    const float inv_det = 1.0f / det;
    const vec3 t = r.o - c;
    const float u = dot(t, p) * inv_det;
    if (u < 0.0f || u > 1.0f) return 0.0f;
    const vec3 q = cross(t, e1);
    const float v = dot(r.d, q) * inv_det;
    if (v < 0.0f || (u + v) > 1.0f) return 0.0f;
    const float dist = dot(e2, q) * inv_det;
    return positive(dist) ? dist : 0.0f;
}

// Sphere::intersect_world distance (src/geom/sphere.rs:70-101).
__device__ __forceinline__ float hit_sphere(const GpuComp& g, const RayFrame& r) {
    const vec3 origin = r.o - mk(g.q0.x, g.q0.y, g.q0.z);
    const float radius = g.q0.w;
    const float a = dot(r.d, r.d);
    const float b = dot(r.d, origin);
    const float c = dot(origin, origin) - (radius * radius);
    const float disc = (b * b) - (a * c);
    if (disc > 0.0f) {
        const float sq = sqrtf(disc);
        const float res_pos = -b + sq;
        const float res_neg = -b - sq;
        if (positive(res_neg)) return res_neg;
        if (positive(res_pos)) return res_pos;
    }
    return 0.0f;
}

__device__ __forceinline__ GpuComp load_comp(const GpuComp* comps, uint32_t slot) {
    const float4* p = reinterpret_cast<const float4*>(comps + slot);
    GpuComp g;
    g.q0 = p[0];
    g.q1 = p[1];
    g.q2 = p[2];
    return g;
}

__device__ __forceinline__ GpuNode load_node(const GpuNode* nodes, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    GpuNode n;
    n.q0 = p[0];
    n.q1 = p[1];
    n.q2 = p[2];
    const float4 w = p[3];
    n.q3 = make_uint4(__float_as_uint(w.x), __float_as_uint(w.y), __float_as_uint(w.z),
                      __float_as_uint(w.w));
    return n;
}

struct TraceOut {
    float t;
    uint32_t slot;      // closest: winning slot; occlusion: != kNone when occluded
    bool overflow;
};

// One traversal.  MODE kClosest: closest hit.  kOcclusionExact: reference
// order, stop at the first accepted hit closer than `limit`.  kOcclusionAnyHit:
// start with closest = limit, which also prunes boxes beyond the target.
template <int MODE, bool COUNT>
__device__ __forceinline__ TraceOut trace(const SceneView& sv, const RayFrame& r, float limit,
                                          uint32_t* cnt) {
    TraceOut out{0.0f, kNone, false};
    float closest = MODE == kOcclusionAnyHit ? limit : FLT_MAX;
    float tmin;
    if (COUNT) cnt[0] += 1;
    if (!(slab(sv.root_lo[0], sv.root_lo[1], sv.root_lo[2], sv.root_hi[0], sv.root_hi[1],
               sv.root_hi[2], r, sv.pad, &tmin) &&
          tmin < closest))
        return out;
    uint32_t stack_ref[kStackSize];
    float stack_t[kStackSize];
    int sp = 0;
    uint32_t cur = sv.root_ref;
    for (;;) {
        if (cur & kLeafBit) {
            uint32_t slot = cur & ~kLeafBit;
            for (;;) {
                const GpuComp g = load_comp(sv.comps, slot);
                const uint32_t tag = __float_as_uint(g.q2.y);
                float dist;
                if (tag & kTagSphere) {
                    if (COUNT) cnt[2] += 1;
                    dist = hit_sphere(g, r);
                } else {
                    if (COUNT) cnt[1] += 1;
                    dist = hit_triangle(g, r);
                }
                if (dist != 0.0f && dist < closest) {
                    closest = dist;
                    out.slot = slot;
                    if (MODE == kOcclusionAnyHit || (MODE == kOcclusionExact && dist < limit)) {
                        out.t = dist;
                        return out;
                    }
                }
                if (tag & kTagLast) break;
                ++slot;
            }
        } else {
            const GpuNode nd = load_node(sv.nodes, cur);
            if (COUNT) cnt[0] += 2;
            float t0 = 0.0f, t1 = 0.0f;
            const bool s0 = slab(nd.q0.x, nd.q0.y, nd.q0.z, nd.q0.w, nd.q1.x, nd.q1.y, r, sv.pad, &t0);
            const bool s1 = slab(nd.q1.z, nd.q1.w, nd.q2.x, nd.q2.y, nd.q2.z, nd.q2.w, r, sv.pad, &t1);
            const uint32_t axis = nd.q3.z;
            const bool flip = axis == 0 ? r.nx : (axis == 1 ? r.ny : r.nz);
            const uint32_t near_ref = flip ? nd.q3.y : nd.q3.x;
            const uint32_t far_ref = flip ? nd.q3.x : nd.q3.y;
            const bool near_ok = flip ? s1 : s0;
            const bool far_ok = flip ? s0 : s1;
            const float near_t = flip ? t1 : t0;
            const float far_t = flip ? t0 : t1;
            if (far_ok) {
                if (sp == kStackSize) {
                    out.overflow = true;
                    out.slot = kNone;
                    return out;
                }
                stack_ref[sp] = far_ref;
                stack_t[sp] = far_t;
                ++sp;
            }
            if (near_ok && near_t < closest) {
                cur = near_ref;
                continue;
            }
        }
        bool found = false;
        while (sp > 0) {
            --sp;
            if (stack_t[sp] < closest) {
                cur = stack_ref[sp];
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    out.t = closest;
    if (MODE != kClosest) out.slot = kNone;  // traversal ended without proof of occlusion
    return out;
}

template <bool COUNT>
__global__ __launch_bounds__(256) void k_intersect(SceneView sv, const gr_ray* __restrict__ rays,
                                                   uint32_t n, gr_hit* __restrict__ hits,
                                                   uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* rp = reinterpret_cast<const float4*>(rays) + 2 * (size_t)i;
    const float4 a = rp[0], b = rp[1];
    const RayFrame r = make_frame(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z));
    uint32_t cnt[3] = {0, 0, 0};
    const TraceOut t = trace<kClosest, COUNT>(sv, r, 0.0f, cnt);
    uint4 h = make_uint4(kNone, kNone, 0u, kNone);
    if (t.overflow) {
        h.z = __float_as_uint(-1.0f);
        atomicAdd(sv.overflow, 1ull);
    } else if (t.slot != kNone) {
        const uint2 id = sv.comp_ids[t.slot];
        h = make_uint4(id.x, id.y, __float_as_uint(t.t), t.slot);
    }
    reinterpret_cast<uint4*>(hits)[i] = h;
    if (COUNT) {
        counts[3 * (size_t)i + 0] = cnt[0];
        counts[3 * (size_t)i + 1] = cnt[1];
        counts[3 * (size_t)i + 2] = cnt[2];
    }
}

// Bvh::visibility, src/geom/bvh.rs:460-477.
template <int MODE, bool COUNT>
__global__ __launch_bounds__(256) void k_occluded(SceneView sv, const gr_segment* __restrict__ segs,
                                                  uint32_t n, uint32_t* __restrict__ visible,
                                                  uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* sp = reinterpret_cast<const float4*>(segs) + 2 * (size_t)i;
    const float4 a = sp[0], b = sp[1];
    const vec3 start = mk(a.x, a.y, a.z), target = mk(b.x, b.y, b.z);
    uint32_t cnt[3] = {0, 0, 0};
    uint32_t vis = 0;
    const bool close = fabsf(target.x - start.x) < 1e-3f && fabsf(target.y - start.y) < 1e-3f &&
                       fabsf(target.z - start.z) < 1e-3f;
    if (!close) {
        const vec3 dir = normalized(target - start);
        const vec3 origin = start + dir * kRayPush;  // Ray::nudge, src/core/ray.rs:44-46
        const float limit = length(target - origin) - 1e-3f;
        const RayFrame r = make_frame(origin, dir);
        const TraceOut t = trace<MODE, COUNT>(sv, r, limit, cnt);
        if (t.overflow) atomicAdd(sv.overflow, 1ull);
        vis = (t.slot == kNone && !t.overflow) ? 1u : 0u;
    }
    visible[i] = vis;
    if (COUNT) {
        counts[3 * (size_t)i + 0] = cnt[0];
        counts[3 * (size_t)i + 1] = cnt[1];
        counts[3 * (size_t)i + 2] = cnt[2];
    }
}

inline dim3 grid_for(uint32_t n) { return dim3((n + 255) / 256); }

}  // namespace

hipError_t launch_intersect(const SceneView& sv, const gr_ray* rays, uint32_t n, gr_hit* hits,
                            uint32_t* counts, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (counts)
        hipLaunchKernelGGL(k_intersect<true>, grid_for(n), dim3(256), 0, s, sv, rays, n, hits, counts);
    else
        hipLaunchKernelGGL(k_intersect<false>, grid_for(n), dim3(256), 0, s, sv, rays, n, hits, counts);
    return hipGetLastError();
}

hipError_t launch_occluded(const SceneView& sv, const gr_segment* segs, uint32_t n,
                           uint32_t* visible, int mode, uint32_t* counts, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 g = grid_for(n), b(256);
    if (mode == GR_OCCLUSION_ANYHIT) {
        if (counts) hipLaunchKernelGGL((k_occluded<kOcclusionAnyHit, true>), g, b, 0, s, sv, segs, n, visible, counts);
        else hipLaunchKernelGGL((k_occluded<kOcclusionAnyHit, false>), g, b, 0, s, sv, segs, n, visible, counts);
    } else {
        if (counts) hipLaunchKernelGGL((k_occluded<kOcclusionExact, true>), g, b, 0, s, sv, segs, n, visible, counts);
        else hipLaunchKernelGGL((k_occluded<kOcclusionExact, false>), g, b, 0, s, sv, segs, n, visible, counts);
    }
    return hipGetLastError();
}

}  // namespace gr
