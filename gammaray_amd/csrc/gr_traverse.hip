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
    bool plain;     // origin and 1/d finite, 1/d nonzero on every axis
    uint32_t sign;  // nx | ny << 1 | nz << 2: dir_is_neg by axis
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
    r.sign = (r.nx ? 1u : 0u) | (r.ny ? 2u : 0u) | (r.nz ? 4u : 0u);
    r.plain = isfinite(o.x) && isfinite(o.y) && isfinite(o.z) && isfinite(r.inv.x) && isfinite(r.inv.y) &&
              isfinite(r.inv.z) && r.inv.x != 0.0f && r.inv.y != 0.0f && r.inv.z != 0.0f;
    return r;
}

// The part of BBox::intersect that does not depend on the current closest
// distance; *tmin receives the entry distance the final `t_min < max_dist`
// compares.  The reference's early returns become one combined condition
// over the same comparisons of the same values (no divergent branches).
__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                     const RayFrame& r, float pad, float* tmin) {
    float t_min = ((r.nx ? hx : lx) - r.o.x) * r.inv.x;
    float t_max = ((r.nx ? lx : hx) - r.o.x) * r.inv.x;
    const float ty_min = ((r.ny ? hy : ly) - r.o.y) * r.inv.y;
    float ty_max = ((r.ny ? ly : hy) - r.o.y) * r.inv.y;
    const float tz_min = ((r.nz ? hz : lz) - r.o.z) * r.inv.z;
    float tz_max = ((r.nz ? lz : hz) - r.o.z) * r.inv.z;
    t_max *= pad;
    ty_max *= pad;
    tz_max *= pad;
    const bool ok_y = !(t_min > ty_max || ty_min > t_max);
    t_min = ty_min > t_min ? ty_min : t_min;
    t_max = ty_max < t_max ? ty_max : t_max;
    const bool ok_z = !(t_min > tz_max || tz_min > t_max);
    t_max = tz_max < t_max ? tz_max : t_max;
    *tmin = t_min;
    return ok_y & ok_z & (t_max > 0.0f);
}

// slab() for a plain ray (RayFrame::plain) and a box that is finite with
// lo <= hi on each axis.  Then no operand is NaN and, per axis, the entry
// distance is <= the exit distance, so the reference's staged comparisons
// reduce to one interval test:
//   pass  <=>  max(tx_min, ty_min, tz_min) <= E  and  E > 0,
//   E = min(tx_max, ty_max, tz_max) * pad,
// with the same entry distance max(tx_min, ty_min) (the z slab never raises
// it).  Multiplying by pad > 0 is monotonic under rounding, so padding the
// minimum equals the minimum of the padded exits bit for bit; a zero's sign
// can differ from the staged form, which only feeds comparisons.
__device__ __forceinline__ bool slab_plain(float lx, float ly, float lz, float hx, float hy, float hz,
                                           const RayFrame& r, float pad, float* tmin) {
    const float tx_min = ((r.nx ? hx : lx) - r.o.x) * r.inv.x;
    const float tx_max = ((r.nx ? lx : hx) - r.o.x) * r.inv.x;
    const float ty_min = ((r.ny ? hy : ly) - r.o.y) * r.inv.y;
    const float ty_max = ((r.ny ? ly : hy) - r.o.y) * r.inv.y;
    const float tz_min = ((r.nz ? hz : lz) - r.o.z) * r.inv.z;
    const float tz_max = ((r.nz ? lz : hz) - r.o.z) * r.inv.z;
    const float entry = __builtin_fmaxf(tx_min, ty_min);
    const float exit = __builtin_fminf(__builtin_fminf(tx_max, ty_max), tz_max) * pad;
    *tmin = entry;
    return (__builtin_fmaxf(entry, tz_min) <= exit) & (exit > 0.0f);
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

// ---------------------------------------------------------------------------
// Persistent traversal.  One 64-lane wave per block; every lane owns one ray
// at a time.  Lanes whose ray finishes go idle and the wave refills them, once
// `refill` are idle, from ray positions it reserved in chunks on a global work
// counter, so long and short rays share a wave without leaving lanes empty.
// Rays are taken in the order of `perm` (a coherence sort) when given; results
// always go to the ray's own index.  Each pass of the loop advances every
// active lane by one step, an interior node or one primitive of a leaf,
// through one shared fetch.  The traversal stack keeps its top LDS entries per
// lane in LDS and spills older entries to private memory, up to the
// reference's 64 entries.
constexpr int kWave = 64;

template <int MODE>
struct Output {
    using type = uint32_t;  // visibility flag
};
template <>
struct Output<kClosest> {
    using type = uint4;  // gr_hit
};

__device__ __forceinline__ uint64_t lanes_below() {
    const uint32_t lane = threadIdx.x;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Eight waves per SIMD keep the most rays in flight (the kernel waits on
// dependent node fetches); instances with deeper LDS stacks are LDS-limited
// below that anyway.
template <int MODE, bool COUNT, int LDS, bool MULTI>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_trace(SceneView sv, const SceneRef* __restrict__ scenes,
                                                  uint32_t nscenes, const float4* __restrict__ in, uint32_t n,
                                                  const uint32_t* __restrict__ perm,
                                                  typename Output<MODE>::type* __restrict__ out,
                                                  uint32_t* __restrict__ counts, uint32_t* __restrict__ next_ray,
                                                  const uint32_t* __restrict__ n_ptr, uint32_t refill,
                                                  uint32_t parts, uint32_t chunk) {
    constexpr int kSpill = kStackSize - LDS;
    // stack entries: (node ref, entry distance bits), one 8-B word per entry
    __shared__ uint2 s_e[LDS][kWave];
    uint2 spill[kSpill];
    const uint32_t lane = threadIdx.x;
    // A ray count that earlier kernels on the stream produced (queue length).
    if (n_ptr) {
        const uint32_t m = __builtin_amdgcn_readfirstlane(*n_ptr);
        n = m < n ? m : n;
    }
    // The grid is sized for the queue's capacity; when fewer rays are queued,
    // waves beyond one per 64 rays have nothing to take.  They leave at once
    // instead of each draining every range's work counter with an atomic.
    if ((uint64_t)blockIdx.x * kWave >= n) return;
    // A chunk no larger than a quarter of each wave's even share, so that a
    // short batch still spreads over every wave of the grid.
    {
        const uint32_t share = n / (gridDim.x * 4u);
        const uint32_t c = share > kWave ? share : (uint32_t)kWave;
        chunk = c < chunk ? c : chunk;
    }

    // With parts > 1 the batch is cut into `parts` contiguous ranges, one per
    // XCD (workgroups are dealt to the XCDs round-robin); a wave drains its own
    // XCD's range first, then helps with the following ones.
    uint32_t part = blockIdx.x % parts;
    uint32_t drained = 0;
    // the current range [r_lo, r_lo + r_len), wave-uniform; recomputed only when
    // the wave moves on to the next range (a 64-bit division when parts > 1)
    uint32_t r_lo = 0, r_len = n;
    auto set_range = [&]() {
        if (parts > 1) {
            r_lo = __builtin_amdgcn_readfirstlane((uint32_t)(((uint64_t)n * part) / parts));
            r_len = __builtin_amdgcn_readfirstlane((uint32_t)(((uint64_t)n * (part + 1)) / parts)) - r_lo;
        }
    };
    set_range();
    bool active = false;
    bool exhausted = false;
    uint32_t ray = 0, cur = 0, best = kNone;
    float closest = FLT_MAX, limit = 0.0f;
    int sp = 0, base = 0;
    uint32_t c_box = 0, c_tri = 0, c_sph = 0;
    RayFrame r;
    r.o = r.d = r.inv = mk(0.0f, 0.0f, 0.0f);
    r.nx = r.ny = r.nz = r.plain = false;
    r.sign = 0u;
    // the lane's scene (the launch's scene unless MULTI)
    uint32_t scene_plain = sv.simple_boxes;
    bool plain = false;  // slab_plain() applies to this lane's ray and scene
    const GpuNode* nodes = sv.nodes;
    const GpuComp* comps = sv.comps;
    const uint2* comp_ids = sv.comp_ids;

    auto finish = [&](bool occluded, bool overflow) {
        if (MODE == kClosest) {
            uint4 h = make_uint4(kNone, kNone, 0u, kNone);
            if (overflow) {
                h.z = __float_as_uint(-1.0f);
            } else if (best != kNone) {
                const uint2 id = comp_ids[best];
                h = make_uint4(id.x, id.y, __float_as_uint(closest), best);
            }
            reinterpret_cast<uint4*>(out)[ray] = h;
        } else {
            reinterpret_cast<uint32_t*>(out)[ray] = (occluded || overflow) ? 0u : 1u;
        }
        if (overflow) atomicAdd(sv.overflow, 1ull);
        if (COUNT) {
            counts[3 * (size_t)ray + 0] = c_box;
            counts[3 * (size_t)ray + 1] = c_tri;
            counts[3 * (size_t)ray + 2] = c_sph;
        }
        active = false;
    };

    // Pops entries until one whose entry distance is still below `closest`.
    // The visibility kernels' counting variant counts each popped entry's box
    // test there (the reference's deferred test of a far child).
    auto pop = [&]() -> bool {
        while (sp > 0) {
            --sp;
            if (COUNT && MODE != kClosest) c_box += 1;
            uint2 e;
            if (sp < base) {
                e = spill[sp];
                base = sp;
            } else {
                e = s_e[sp & (LDS - 1)][lane];
            }
            if (__uint_as_float(e.y) < closest) {
                cur = e.x;
                return true;
            }
        }
        return false;
    };

    // Wave-local pool of reserved ray positions [pool_lo + pool_k, pool_lo +
    // pool_end).  A wave reserves `chunk` positions per atomic on its range's
    // work counter and refills from the pool: one counter serves every wave of
    // the launch, and atomics on one address serialise in the L2.
    uint32_t pool_k = 0, pool_end = 0, pool_lo = 0;

    for (;;) {
        // ---- refill idle lanes (wave-uniform control flow)
        if (!exhausted || pool_k < pool_end) {
            const uint64_t idle = __ballot(!active);
            const uint32_t nidle = (uint32_t)__popcll(idle);
            if (nidle >= refill) {
                const uint32_t avail = pool_end - pool_k;
                uint32_t nfirst = 0, nsize = 0, nlo = 0;
                if (avail < nidle && !exhausted) {
                    const uint32_t lo = r_lo, len = r_len;
                    uint32_t first = 0;
                    if (lane == 0) first = atomicAdd(next_ray + part * kCounterStride, chunk);
                    first = __builtin_amdgcn_readfirstlane(first);
                    if (first < len) {
                        nfirst = first;
                        nsize = len - first < chunk ? len - first : chunk;
                        nlo = lo;
                    }
                    if ((uint64_t)first + chunk >= len) {
                        // this range is drained (counters only grow): go to the next
                        if (++drained == parts) exhausted = true;
                        part = part + 1 == parts ? 0u : part + 1;
                        set_range();
                    }
                }
                if (!active) {
                    // idle lanes take the pool's positions first, then the new chunk's
                    const uint32_t rank = (uint32_t)__popcll(idle & lanes_below());
                    uint32_t idx = n;
                    if (rank < avail) idx = pool_lo + pool_k + rank;
                    else if (rank - avail < nsize) idx = nlo + nfirst + (rank - avail);
                    // A permutation entry outside [0, n) is skipped (never
                    // dereferenced), so a faulty sort shows as missing results.
                    const uint32_t pick = idx < n ? (perm ? perm[idx] : idx) : n;
                    if (pick < n) {
                        ray = pick;
                        const float4 a = in[2 * (size_t)ray], b = in[2 * (size_t)ray + 1];
                        best = kNone;
                        sp = 0;
                        base = 0;
                        c_box = c_tri = c_sph = 0;
                        bool start = true;
                        if (MODE == kClosest) {
                            r = make_frame(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z));
                            closest = FLT_MAX;
                        } else {
                            // Bvh::visibility, src/geom/bvh.rs:460-477.
                            const vec3 s0 = mk(a.x, a.y, a.z), tg = mk(b.x, b.y, b.z);
                            const bool close = fabsf(tg.x - s0.x) < 1e-3f && fabsf(tg.y - s0.y) < 1e-3f &&
                                               fabsf(tg.z - s0.z) < 1e-3f;
                            if (close) {
                                start = false;
                                finish(true, false);
                            } else {
                                const vec3 dir = normalized(tg - s0);
                                const vec3 org = s0 + dir * kRayPush;  // Ray::nudge, ray.rs:44-46
                                limit = length(tg - org) - 1e-3f;
                                r = make_frame(org, dir);
                                closest = MODE == kOcclusionAnyHit ? limit : FLT_MAX;
                            }
                        }
                        if (start) {
                            float tmin;
                            if (COUNT) c_box = 1;
                            uint32_t root = sv.root_ref;
                            float lo0 = sv.root_lo[0], lo1 = sv.root_lo[1], lo2 = sv.root_lo[2];
                            float hi0 = sv.root_hi[0], hi1 = sv.root_hi[1], hi2 = sv.root_hi[2];
                            if (MULTI) {
                                const uint32_t si = __float_as_uint(a.w);
                                const SceneRef& sr = scenes[si < nscenes ? si : 0u];
                                nodes = sr.nodes;
                                comps = sr.comps;
                                comp_ids = sr.comp_ids;
                                root = sr.root_ref;
                                scene_plain = sr.simple_boxes;
                                lo0 = sr.root_lo[0], lo1 = sr.root_lo[1], lo2 = sr.root_lo[2];
                                hi0 = sr.root_hi[0], hi1 = sr.root_hi[1], hi2 = sr.root_hi[2];
                            }
                            plain = r.plain && scene_plain != 0u;
                            if (slab(lo0, lo1, lo2, hi0, hi1, hi2, r, sv.pad, &tmin) && tmin < closest) {
                                cur = root;
                                active = true;
                            } else {
                                finish(false, false);
                            }
                        }
                    }
                }
                if (nidle <= avail) {
                    pool_k += nidle;
                } else if (nsize) {
                    const uint32_t used = nidle - avail < nsize ? nidle - avail : nsize;
                    pool_lo = nlo;
                    pool_k = nfirst + used;
                    pool_end = nfirst + nsize;
                } else {
                    pool_k = pool_end;
                }
            }
        }
        if (__ballot(active) == 0) {
            if (exhausted && pool_k >= pool_end) break;
            continue;
        }

        // ---- one step for every active lane: an interior node, or one
        // primitive of its leaf.  Both fetch through the same four loads (a
        // primitive record is 48 bytes; its fourth load repeats the first), so
        // a wave whose lanes are at different kinds of nodes still waits on one
        // round trip per step.
        if (active) {
            const bool at_leaf = (cur & kLeafBit) != 0;
            const uint32_t slot = cur & ~kLeafBit;
            const float4* p = at_leaf ? reinterpret_cast<const float4*>(comps + slot)
                                      : reinterpret_cast<const float4*>(nodes + cur);
            const float4 q0 = p[0], q1 = p[1], q2 = p[2];
            const float4 q3 = *(at_leaf ? p : p + 3);
            // All four loads complete here, before the interior/leaf branch:
            // left to itself the compiler sinks the loads only the interior
            // path reads into that path, which costs a wave holding both
            // kinds of lanes a second round trip after the leaf path's wait.
            asm volatile("" ::"v"(q0.x), "v"(q0.y), "v"(q0.z), "v"(q0.w), "v"(q1.x), "v"(q1.y), "v"(q1.z),
                         "v"(q1.w), "v"(q2.x), "v"(q2.y), "v"(q2.z), "v"(q2.w), "v"(q3.x), "v"(q3.y),
                         "v"(q3.z), "v"(q3.w));
            if (!at_leaf) {
                // Closest hit counts both child boxes here; visibility counts the
                // near child's here and the far child's at its pop (pop()).
                if (COUNT) c_box += MODE == kClosest ? 2 : 1;
                float t0 = 0.0f, t1 = 0.0f;
                bool s0, s1;
                if (plain) {
                    s0 = slab_plain(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, r, sv.pad, &t0);
                    s1 = slab_plain(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, r, sv.pad, &t1);
                } else {
                    s0 = slab(q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, r, sv.pad, &t0);
                    s1 = slab(q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, r, sv.pad, &t1);
                }
                const uint32_t axis = __float_as_uint(q3.z);
                const bool flip = ((r.sign >> axis) & 1u) != 0u;
                const uint32_t near_ref = __float_as_uint(flip ? q3.y : q3.x);
                const uint32_t far_ref = __float_as_uint(flip ? q3.x : q3.y);
                const bool near_ok = flip ? s1 : s0;
                const bool far_ok = flip ? s0 : s1;
                const float near_t = flip ? t1 : t0;
                const float far_t = flip ? t0 : t1;
                // A far child whose entry distance is already >= closest would fail
                // the deferred test at its pop (closest only decreases): not pushed.
                // The visibility kernels' counting variant pushes every far child,
                // as the reference does, so that a traversal ending at its first
                // occluder counts exactly the boxes the reference tests before it
                // stops (those still on the stack then are never tested); a child
                // whose slab test failed carries an infinite entry distance and
                // fails at its pop.
                const bool push = (COUNT && MODE != kClosest) || (far_ok && far_t < closest);
                if (push && sp == kStackSize) {
                    finish(false, true);
                } else {
                    if (push) {
                        if (sp - base == LDS) {
                            spill[base] = s_e[base & (LDS - 1)][lane];
                            ++base;
                        }
                        s_e[sp & (LDS - 1)][lane] =
                            make_uint2(far_ref, __float_as_uint(far_ok ? far_t : __builtin_inff()));
                        ++sp;
                    }
                    if (near_ok && near_t < closest) {
                        cur = near_ref;
                    } else if (!pop()) {
                        finish(false, false);
                    }
                }
            } else {
                GpuComp g;
                g.q0 = q0;
                g.q1 = q1;
                g.q2 = q2;
                const uint32_t tag = __float_as_uint(g.q2.y);
                float dist;
                if (tag & kTagSphere) {
                    if (COUNT) c_sph += 1;
                    dist = hit_sphere(g, r);
                } else {
                    if (COUNT) c_tri += 1;
                    dist = hit_triangle(g, r);
                }
                bool done = false;
                if (dist != 0.0f && dist < closest) {
                    closest = dist;
                    best = slot;
                    done = MODE == kOcclusionAnyHit || (MODE == kOcclusionExact && dist < limit);
                }
                if (done) finish(true, false);
                else if (!(tag & kTagLast)) cur = kLeafBit | (slot + 1);
                else if (!pop()) finish(false, false);
            }
        }
    }
}

__global__ void k_reset_counter(uint32_t* c) {
    for (uint32_t i = 0; i < kRayCounters; ++i) c[i * kCounterStride] = 0u;
}

template <int MODE, bool COUNT, int LDS, bool MULTI>
int occupancy() {
    static int per_cu = -1;
    if (per_cu < 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_trace<MODE, COUNT, LDS, MULTI>, kWave, 0) != hipSuccess)
            v = 1;
        per_cu = v > 0 ? v : 1;
    }
    return per_cu;
}

int num_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            cus = v;
        else
            return 1;
    }
    return cus;
}

template <int MODE, bool COUNT, int LDS, bool MULTI>
hipError_t launch_lds(const TraceSetup& ts, const void* in, uint32_t n, const uint32_t* n_ptr, const uint32_t* perm,
                      void* out, uint32_t* counts, hipStream_t s) {
    int per_cu = occupancy<MODE, COUNT, LDS, MULTI>();
    if (ts.tu.waves_per_cu > 0 && (int)ts.tu.waves_per_cu < per_cu) per_cu = (int)ts.tu.waves_per_cu;
    const uint32_t grid = (uint32_t)(num_cus() * per_cu);
    const uint32_t needed = (uint32_t)(((uint64_t)n + kWave - 1) / kWave);
    const uint32_t blocks = needed < grid ? needed : grid;
    const uint32_t want = MODE == kClosest ? ts.tu.refill : ts.tu.vis_refill;
    const uint32_t refill = want < 1 ? 1u : (want > 64 ? 64u : want);
    const uint32_t parts = ts.tu.xcd_ranges ? kRayCounters : 1u;
    // A range's 32-bit work counter ends below len + (waves x chunk): each wave
    // adds at most one chunk past the end before it moves on, so the chunk is
    // capped to keep that below 2^32 for the largest batches.
    uint32_t chunk = ts.tu.ray_chunk < 1 ? 1u : ts.tu.ray_chunk;
    const uint64_t room = blocks ? (0xffffffffull - n) / blocks : 0xffffffffull;
    if (chunk > room) chunk = room > 0 ? (uint32_t)room : 1u;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, ts.sv.next_ray);
    hipEvent_t t0 = ts.timing ? timing_begin(ts.timing, s) : nullptr;
    hipLaunchKernelGGL((k_trace<MODE, COUNT, LDS, MULTI>), dim3(blocks), dim3(kWave), 0, s, ts.sv, ts.scenes,
                       ts.nscenes,
                       reinterpret_cast<const float4*>(in), n, perm,
                       reinterpret_cast<typename Output<MODE>::type*>(out), counts, ts.sv.next_ray, n_ptr, refill,
                       parts, chunk);
    if (t0) timing_end(ts.timing, MODE == kClosest ? GR_KERNEL_TRACE_CLOSEST : GR_KERNEL_TRACE_SHADOW, t0, s);
    return hipGetLastError();
}

template <int MODE, bool COUNT, bool MULTI>
hipError_t launch_sorted(const TraceSetup& ts, const void* in, uint32_t n, const uint32_t* n_ptr, void* out,
                         uint32_t* counts, hipStream_t s) {
    const uint32_t* perm = nullptr;
    if (ts.tu.sort_bits != 0) {
        if (!ts.sort) return hipErrorInvalidValue;
        hipEvent_t t0 = ts.timing ? timing_begin(ts.timing, s) : nullptr;
        const hipError_t e =
            MULTI ? sort_rays_multi(ts.scenes, ts.nscenes, reinterpret_cast<const float4*>(in), n, n_ptr,
                                    MODE != kClosest, ts.tu.sort_bits, *ts.sort, &perm, s)
                  : sort_rays(ts.sv, reinterpret_cast<const float4*>(in), n, n_ptr, MODE != kClosest,
                              ts.tu.sort_bits, *ts.sort, &perm, s);
        if (t0) timing_end(ts.timing, GR_KERNEL_RAY_SORT, t0, s);
        if (e != hipSuccess) return e;
    }
    switch (ts.tu.lds_stack) {
        case 8: return launch_lds<MODE, COUNT, 8, MULTI>(ts, in, n, n_ptr, perm, out, counts, s);
        case 32: return launch_lds<MODE, COUNT, 32, MULTI>(ts, in, n, n_ptr, perm, out, counts, s);
        default: return launch_lds<MODE, COUNT, 16, MULTI>(ts, in, n, n_ptr, perm, out, counts, s);
    }
}

template <int MODE, bool COUNT>
hipError_t launch_trace(const TraceSetup& ts, const void* in, uint32_t n, const uint32_t* n_ptr, void* out,
                        uint32_t* counts, hipStream_t s) {
    return ts.scenes ? launch_sorted<MODE, COUNT, true>(ts, in, n, n_ptr, out, counts, s)
                     : launch_sorted<MODE, COUNT, false>(ts, in, n, n_ptr, out, counts, s);
}

}  // namespace

int trace_occupancy(uint32_t lds_stack) {
    switch (lds_stack) {
        case 8: return occupancy<kClosest, false, 8, false>();
        case 32: return occupancy<kClosest, false, 32, false>();
        default: return occupancy<kClosest, false, 16, false>();
    }
}

hipError_t launch_intersect(const TraceSetup& ts, const gr_ray* rays, uint32_t n, const uint32_t* n_ptr,
                            gr_hit* hits, uint32_t* counts, hipStream_t s) {
    if (n == 0) return hipSuccess;
    return counts ? launch_trace<kClosest, true>(ts, rays, n, n_ptr, hits, counts, s)
                  : launch_trace<kClosest, false>(ts, rays, n, n_ptr, hits, counts, s);
}

hipError_t launch_occluded(const TraceSetup& ts, const gr_segment* segs, uint32_t n, const uint32_t* n_ptr,
                           uint32_t* visible, int mode, uint32_t* counts, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (mode == GR_OCCLUSION_ANYHIT)
        return counts ? launch_trace<kOcclusionAnyHit, true>(ts, segs, n, n_ptr, visible, counts, s)
                      : launch_trace<kOcclusionAnyHit, false>(ts, segs, n, n_ptr, visible, counts, s);
    return counts ? launch_trace<kOcclusionExact, true>(ts, segs, n, n_ptr, visible, counts, s)
                  : launch_trace<kOcclusionExact, false>(ts, segs, n, n_ptr, visible, counts, s);
}

}  // namespace gr
