// gr_bvh.cpp -- host BVH construction with the reference's SAH rules, and the
// conversion of the resulting tree into the 64-byte GPU node layout.
//
// The tree must be identical to the one Bvh::build produces
// (src/geom/bvh.rs:182-382), because traversal results are checked bit for bit
// and ties are resolved by visiting order.  The builder below is iterative: it
// emits nodes directly in depth-first preorder (the order flatten_tree writes,
// bvh.rs:310-337) from an explicit task stack, left subtree first, so the
// ordered component list comes out in the same leaf order as the recursion.
// The split rules it applies per node are the reference's:
//   * one component -> leaf;
//   * centroid bbox degenerate along its maximum_extent axis -> one leaf holding
//     every component (bvh.rs:207-216);
//   * n <= 4 -> median split with the reference's nth_element (util.rs:23-46),
//     whose left recursion skips one element;
//   * otherwise 12-bucket SAH with the bbox "surface area" that is really a
//     volume (bbox.rs:73-76), split when n > 255 or min_cost < n, else a median
//     split (bvh.rs:229-296).
#pragma clang fp contract(off)

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <utility>

#include "gr_internal.h"

namespace gr {
namespace {

constexpr int kBuckets = 12;

struct Box {
    float lo[3], hi[3];
};

inline Box empty_box() {
    // BBox::empty, src/core/bbox.rs:15-20 (f32::MIN is -FLT_MAX).
    return Box{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
}

// Accumulator first, as combine_with/union_with evaluate f32::min(self, other).
inline void grow(Box& b, const float lo[3], const float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = fminf(b.lo[k], lo[k]);
        b.hi[k] = fmaxf(b.hi[k], hi[k]);
    }
}

inline int widest_axis(const Box& b) {
    // maximum_extent, src/core/bbox.rs:52-63 (ties resolve as written there).
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (dx >= dy && dy >= dz) return 0;
    if (dy >= dz) return 1;
    return 2;
}

inline float volume(const Box& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy * dz;
}

struct Item {
    float lo[3], hi[3], c[3];
    uint32_t prim, comp, bucket;
};

// (12.0 * rel[dim]) as usize, clamped to [0, 11]; Rust's cast saturates and
// maps NaN to 0.
inline uint32_t bucket_index(const Box& cb, const float c[3], int dim) {
    float rel = (c[dim] - cb.lo[dim]) / (cb.hi[dim] - cb.lo[dim]);
    float f = (float)kBuckets * rel;
    uint64_t b;
    if (!(f > 0.0f)) b = 0;
    else if (f >= 18446744073709551616.0f) b = UINT64_MAX;
    else b = (uint64_t)f;
    return b > kBuckets - 1 ? (uint32_t)(kBuckets - 1) : (uint32_t)b;
}

// Quickselect of src/geom/util.rs:23-46: pivot = middle element moved to the
// end, Lomuto partition by strict <, and the left range that stops one short
// of the pivot.  Its exact permutation decides the leaf order, so it is
// reproduced rather than replaced by std::nth_element.
void quickselect(Item* s, size_t len, size_t nth, int dim) {
    while (len >= 2) {
        std::swap(s[len / 2], s[len - 1]);
        const float pv = s[len - 1].c[dim];
        size_t cut = 0;
        for (size_t i = 0; i + 1 < len; ++i) {
            if (s[i].c[dim] < pv) {
                std::swap(s[i], s[cut]);
                ++cut;
            }
        }
        std::swap(s[len - 1], s[cut]);
        if (nth < cut) {
            len = cut - 1;
        } else if (nth > cut) {
            s += cut + 1;
            len -= cut + 1;
            nth -= cut + 1;
        } else {
            return;
        }
    }
}

// Stable-position partition of src/geom/util.rs:4-17 by bucket <= split.
size_t split_by_bucket(Item* s, size_t len, uint32_t split) {
    size_t cut = 0;
    for (size_t i = 0; i < len; ++i) {
        if (s[i].bucket <= split) {
            std::swap(s[i], s[cut]);
            ++cut;
        }
    }
    return cut;
}

// Chooses the split of a node that is not a leaf; returns mid in (0, n).
size_t choose_split(Item* it, size_t n, const Box& bounds, const Box& cb, int dim) {
    if (n <= 4) {
        size_t mid = n / 2;
        quickselect(it, n, mid, dim);
        return mid;
    }
    uint32_t count[kBuckets] = {0};
    Box bb[kBuckets];
    for (int b = 0; b < kBuckets; ++b) bb[b] = empty_box();
    for (size_t i = 0; i < n; ++i) {
        uint32_t b = bucket_index(cb, it[i].c, dim);
        it[i].bucket = b;
        count[b] += 1;
        grow(bb[b], it[i].lo, it[i].hi);
    }
    const float parent_volume = volume(bounds);
    float best_cost = 0.0f;
    uint32_t best = 0;
    for (int s = 0; s < kBuckets - 1; ++s) {
        Box left = empty_box(), right = empty_box();
        size_t nl = 0, nr = 0;
        for (int b = 0; b <= s; ++b) {
            grow(left, bb[b].lo, bb[b].hi);
            nl += count[b];
        }
        for (int b = s + 1; b < kBuckets; ++b) {
            grow(right, bb[b].lo, bb[b].hi);
            nr += count[b];
        }
        float cost = 1.0f + ((float)nl * volume(left) + (float)nr * volume(right)) / parent_volume;
        // First cost seeds the minimum; later ones replace it on strict <
        // (NaN costs never win, bvh.rs:266-273).
        if (s == 0 || cost < best_cost) {
            best_cost = cost;
            best = (uint32_t)s;
        }
    }
    if (n > 255 || best_cost < (float)n) return split_by_bucket(it, n, best);
    size_t mid = n / 2;
    quickselect(it, n, mid, dim);
    return mid;
}

}  // namespace

FlatTree build_reference_bvh(const std::vector<Prim>& prims) {
    FlatTree out;
    std::vector<Item> items;
    size_t total = 0;
    for (const Prim& p : prims) total += p.num_components();
    items.reserve(total);
    for (uint32_t pi = 0; pi < prims.size(); ++pi) {
        const Prim& p = prims[pi];
        for (uint32_t c = 0; c < p.num_components(); ++c) {
            Item it;
            if (p.type == kPrimMesh) {
                // Mesh::bbox_world, src/geom/mesh.rs:262-268.
                Box b = empty_box();
                for (int k = 0; k < 3; ++k) {
                    const vec3& v = p.verts[p.tris[9 * (size_t)c + k]];
                    float pt[3] = {v.x, v.y, v.z};
                    grow(b, pt, pt);
                }
                std::memcpy(it.lo, b.lo, sizeof(b.lo));
                std::memcpy(it.hi, b.hi, sizeof(b.hi));
            } else {
                // Sphere::bbox_world, src/geom/sphere.rs:57-68.
                float cc[3] = {p.center.x, p.center.y, p.center.z};
                for (int k = 0; k < 3; ++k) {
                    it.lo[k] = cc[k] - p.radius;
                    it.hi[k] = cc[k] + p.radius;
                }
            }
            // BvhComponentInfo centroid, src/geom/bvh.rs:26.
            for (int k = 0; k < 3; ++k) it.c[k] = 0.5f * it.lo[k] + 0.5f * it.hi[k];
            it.prim = pi;
            it.comp = c;
            it.bucket = 0;
            items.push_back(it);
        }
    }
    if (items.empty()) return out;

    out.nodes.reserve(2 * items.size());
    out.comp_prim.reserve(items.size());
    out.comp_index.reserve(items.size());

    struct Task {
        size_t begin, end;
        uint32_t parent, depth;
        bool second;
    };
    std::vector<Task> todo;
    todo.push_back(Task{0, items.size(), kNone, 0, false});
    while (!todo.empty()) {
        Task t = todo.back();
        todo.pop_back();
        const uint32_t idx = (uint32_t)out.nodes.size();
        out.nodes.push_back(FlatNode{});
        if (t.second) out.nodes[t.parent].offset = idx;
        out.max_depth = std::max(out.max_depth, t.depth);

        Item* it = items.data() + t.begin;
        const size_t n = t.end - t.begin;
        Box bounds = empty_box();
        for (size_t i = 0; i < n; ++i) grow(bounds, it[i].lo, it[i].hi);
        FlatNode& nd = out.nodes[idx];
        std::memcpy(nd.lo, bounds.lo, sizeof(nd.lo));
        std::memcpy(nd.hi, bounds.hi, sizeof(nd.hi));
        nd.axis = 0;

        bool leaf = n == 1;
        int dim = 0;
        Box cb = empty_box();
        if (!leaf) {
            for (size_t i = 0; i < n; ++i) grow(cb, it[i].c, it[i].c);
            dim = widest_axis(cb);
            leaf = cb.lo[dim] == cb.hi[dim];
        }
        if (leaf) {
            nd.offset = (uint32_t)out.comp_prim.size();
            nd.count = (uint32_t)n;
            for (size_t i = 0; i < n; ++i) {
                out.comp_prim.push_back(it[i].prim);
                out.comp_index.push_back(it[i].comp);
            }
            continue;
        }
        const size_t mid = choose_split(it, n, bounds, cb, dim);
        FlatNode& nd2 = out.nodes[idx];
        nd2.count = 0;
        nd2.axis = (uint32_t)dim;
        todo.push_back(Task{t.begin + mid, t.end, idx, t.depth + 1, true});
        todo.push_back(Task{t.begin, t.begin + mid, idx, t.depth + 1, false});
    }
    // Interior boxes are the union of the two child boxes, as
    // BvhBuildNode::new_interior computes them (bvh.rs:72-84); children have
    // larger preorder indices, so one reverse sweep suffices.
    for (size_t i = out.nodes.size(); i-- > 0;) {
        FlatNode& nd = out.nodes[i];
        if (nd.count != 0) continue;
        const FlatNode& a = out.nodes[i + 1];
        const FlatNode& b = out.nodes[nd.offset];
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = fminf(a.lo[k], b.lo[k]);
            nd.hi[k] = fmaxf(a.hi[k], b.hi[k]);
        }
    }
    return out;
}

GpuLayout make_gpu_layout(const FlatTree& t, const std::vector<Prim>& prims) {
    GpuLayout g;
    const size_t nn = t.nodes.size();
    std::vector<uint32_t> rank(nn, kNone);
    uint32_t ni = 0;
    for (size_t i = 0; i < nn; ++i)
        if (t.nodes[i].count == 0) rank[i] = ni++;
    auto ref = [&](size_t c) -> uint32_t {
        return t.nodes[c].count ? (kLeafBit | t.nodes[c].offset) : rank[c];
    };
    g.nodes.resize(ni);
    g.simple_boxes = true;
    auto simple = [](const FlatNode& c) {
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(c.lo[k]) || !std::isfinite(c.hi[k]) || !(c.lo[k] <= c.hi[k])) return false;
        return true;
    };
    for (size_t i = 0; i < nn; ++i) {
        const FlatNode& nd = t.nodes[i];
        if (nd.count != 0) continue;
        const FlatNode& a = t.nodes[i + 1];
        const FlatNode& b = t.nodes[nd.offset];
        GpuNode& o = g.nodes[rank[i]];
        g.simple_boxes = g.simple_boxes && simple(a) && simple(b);
        o.q0 = make_float4(a.lo[0], a.lo[1], a.lo[2], a.hi[0]);
        o.q1 = make_float4(a.hi[1], a.hi[2], b.lo[0], b.lo[1]);
        o.q2 = make_float4(b.lo[2], b.hi[0], b.hi[1], b.hi[2]);
        o.q3 = make_uint4(ref(i + 1), ref(nd.offset), nd.axis, 0u);
    }
    const size_t nc = t.comp_prim.size();
    g.comps.resize(nc);
    g.comp_ids.resize(nc);
    std::vector<uint32_t> tag(nc, 0);
    for (size_t i = 0; i < nn; ++i)
        if (t.nodes[i].count) tag[t.nodes[i].offset + t.nodes[i].count - 1] |= kTagLast;
    for (size_t k = 0; k < nc; ++k) {
        const Prim& p = prims[t.comp_prim[k]];
        const uint32_t c = t.comp_index[k];
        GpuComp& r = g.comps[k];
        uint32_t tg = tag[k];
        if (p.type == kPrimMesh) {
            const uint32_t* tri = &p.tris[9 * (size_t)c];
            const vec3 A = p.verts[tri[0]], B = p.verts[tri[1]], C = p.verts[tri[2]];
            const vec3 e1 = A - C, e2 = B - C;  // edge1, edge2 of mesh.rs:278-279
            r.q0 = make_float4(C.x, C.y, C.z, e1.x);
            r.q1 = make_float4(e1.y, e1.z, e2.x, e2.y);
            r.q2 = make_float4(e2.z, 0.0f, 0.0f, 0.0f);
        } else {
            tg |= kTagSphere;
            r.q0 = make_float4(p.center.x, p.center.y, p.center.z, p.radius);
            r.q1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            r.q2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        std::memcpy(&r.q2.y, &tg, sizeof(tg));
        g.comp_ids[k] = make_uint2(t.comp_prim[k], c);
    }
    g.root_ref = nn ? ref(0) : kNone;
    for (int k = 0; k < 3; ++k) {
        g.root_lo[k] = nn ? t.nodes[0].lo[k] : 0.0f;
        g.root_hi[k] = nn ? t.nodes[0].hi[k] : 0.0f;
    }
    return g;
}

}  // namespace gr
