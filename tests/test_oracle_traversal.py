"""CPU tests: pin the oracle's BVH build and traversal, and check the
library's host-side builder against it.

What pins what:
  * Philox4x32-10 -- the published Random123 known-answer vectors.
  * BVH build -- structural invariants on several scenes, plus hand-derived
    trees that expose the reference's quirks (median split through the
    nth_element whose left recursion drops an element, src/geom/util.rs:39-41;
    multi-component leaves for coincident centroids, src/geom/bvh.rs:207-216).
    The 12-bucket SAH path (bvh.rs:229-296) is pinned by trees derived by hand
    from its cost arithmetic.  Each case is built so that one plausible
    mistake changes the tree:
      - volume instead of area (Q1): volume and area pick different buckets;
      - the "leaf" branch's median split when min cost >= n;
      - the n > 255 rule forcing a split;
      - all-NaN costs on flat boxes;
      - a NaN cost[0] kept by the strict-< minimum seeded with cost[0].
    The library's builder must give the same trees.
  * intersect -- the reference's own brute force, Bvh::intersect_naive
    (src/geom/bvh.rs:385-400), equal except documented ties (|dt| <= 1e-6 t),
    and closed-form ray/triangle and ray/sphere distances.
  * visibility -- brute-force segment test built on intersect_naive.
"""
import re

import numpy as np
import pytest

from gammaray_amd import scenes
from oracle.oracle import NONE, OracleScene, philox4x32_10


# --------------------------------------------------------------------------
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds.
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(philox4x32_10(ctr, key)) == want


# --------------------------------------------------------------------------
def _spheres_scene(centers, radius=0.1):
    s = scenes.SceneSpec(name="spheres")
    for c in centers:
        s.add_sphere(c, radius, scenes.disney())
    return s


def check_invariants(spec, tree):
    bbox, off, cnt, axis = tree["bbox"], tree["offset"], tree["count"], tree["axis"]
    n = len(off)
    # every (prim, comp) exactly once
    pairs = set(zip(tree["comp_prim"].tolist(), tree["comp_index"].tolist()))
    assert len(pairs) == spec.num_components == len(tree["comp_prim"])
    covered = np.zeros(len(tree["comp_prim"]), np.int32)
    # DFS layout: subtree of i spans [i, end(i)); first child i+1, second = offset
    end = np.zeros(n, np.int64)
    for i in range(n - 1, -1, -1):
        if cnt[i] > 0:
            end[i] = i + 1
            covered[off[i]:off[i] + cnt[i]] += 1
        else:
            assert off[i] > i + 1
            assert end[i + 1] == off[i], "first subtree must end where the second starts"
            end[i] = end[off[i]]
            for c in (i + 1, off[i]):
                assert np.all(bbox[c, :3] >= bbox[i, :3]) and np.all(bbox[c, 3:] <= bbox[i, 3:])
            assert axis[i] in (0, 1, 2)
    assert end[0] == n
    assert np.all(covered == 1)
    # leaf boxes bound their components
    for i in np.nonzero(cnt > 0)[0][:2000]:
        for k in range(off[i], off[i] + cnt[i]):
            p = spec.prims[tree["comp_prim"][k]]
            if p.kind == "mesh":
                v = p.vertices[p.tris[tree["comp_index"][k], :3]]
                lo, hi = v.min(0), v.max(0)
            else:
                lo, hi = p.center - p.radius, p.center + p.radius
            assert np.all(lo >= bbox[i, :3]) and np.all(hi <= bbox[i, 3:])


SCENES = {
    "box": lambda: scenes.box_scene(),
    "ico4": lambda: scenes.icosphere_scene(level=4),
    "showcase_small": lambda: scenes.showcase_scene(n_spheres=40, grid=(30, 20)),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_bvh_invariants(name):
    spec = SCENES[name]()
    check_invariants(spec, OracleScene(spec).export_bvh())


def test_bvh_median_split_quirk():
    # Four spheres on the x axis, centroids 1, 3, 9, 2 (prims 0..3).  n <= 4
    # takes the median split; the reference's nth_element leaves [1, 3, 2, 9]
    # (its left recursion never revisits index 2), so the halves are {1, 3}
    # and {2, 9} -- a correct median split would give {1, 2} | {3, 9}.
    spec = _spheres_scene([(1, 0, 0), (3, 0, 0), (9, 0, 0), (2, 0, 0)])
    t = OracleScene(spec).export_bvh()
    assert t["count"].tolist() == [0, 0, 1, 1, 0, 1, 1]
    assert t["offset"].tolist() == [4, 3, 0, 1, 6, 2, 3]
    assert t["axis"].tolist() == [0, 0, 0, 0, 0, 0, 0]
    assert t["comp_prim"].tolist() == [0, 1, 3, 2]
    np.testing.assert_allclose(t["bbox"][1], [0.9, -0.1, -0.1, 3.1, 0.1, 0.1], rtol=1e-6)


def test_bvh_three_in_a_row():
    spec = _spheres_scene([(0, 0, 0), (1, 0, 0), (2, 0, 0)])
    t = OracleScene(spec).export_bvh()
    assert t["count"].tolist() == [0, 1, 0, 1, 1]
    assert t["offset"].tolist() == [2, 0, 4, 1, 2]
    assert t["comp_prim"].tolist() == [0, 1, 2]


def test_bvh_coincident_centroids_make_one_leaf():
    # Identical spheres: the centroid bbox is a point, so the reference makes
    # a single multi-component leaf (src/geom/bvh.rs:207-216), even above 255.
    spec = _spheres_scene([(0.5, 0.5, 0.5)] * 300)
    t = OracleScene(spec).export_bvh()
    assert t["count"].tolist() == [300]
    assert t["comp_prim"].tolist() == list(range(300))


def test_bvh_axis_choice_prefers_x_on_ties():
    # maximum_extent returns 0 only when dx >= dy and dy >= dz (bbox.rs:52-63):
    # extents (1, 1, 1) -> axis 0; (1, 2, 2) -> axis 1.
    t = OracleScene(_spheres_scene([(0, 0, 0), (1, 1, 1)])).export_bvh()
    assert t["axis"][0] == 0
    t = OracleScene(_spheres_scene([(0, 0, 0), (1, 2, 2)])).export_bvh()
    assert t["axis"][0] == 1


def _subtree_prims(t, node):
    """Prims of the components under `node` (DFS layout: a leaf's range, or
    both children's)."""
    if t["count"][node] > 0:
        o = t["offset"][node]
        return t["comp_prim"][o:o + t["count"][node]].tolist()
    return _subtree_prims(t, node + 1) + _subtree_prims(t, t["offset"][node])


def _volume_spheres():
    spec = _spheres_scene([(12, 0, 0), (4, 0, 0), (0, 0, 0), (5, 0, 0), (1, 0, 0)], radius=1.0)
    for p, r in zip(spec.prims, (3.0, 3.0, 1.0, 3.0, 1.0)):
        p.radius = r
    return spec


def test_bvh_sah_uses_volume():
    """Five spheres on the x axis, given in the order x = 12, 4, 0, 5, 1 (prims
    0..4) with radii 3, 3, 1, 3, 1.  Centroids span 0..12, so bucket =
    floor(12 x / 12) = [11, 4, 0, 5, 1].  The reference's "surface area" is the
    box volume (Q1, src/core/bbox.rs:73-76).  The node box is [-1,15] x [-3,3]^2,
    volume 576.  Split costs 1 + (n0 V0 + n1 V1) / 576 after bucket:
      0:     {0} V 8, {1,4,5,12} V 540      -> 1 + 2168/576 = 4.764
      1..3:  {0,1} V 12, {4,5,12} V 504     -> 1 + 1536/576 = 3.667
      4:     {0,1,4} V 288, {5,12} V 468    -> 1 + 1800/576 = 4.125
      5..10: {0,1,4,5} V 324, {12} V 216    -> 1 + 1512/576 = 3.625  (minimum)
    A true surface area (node 456) would give 3.825 after bucket 1 and 4.0
    after bucket 5, and split {0,1} | {4,5,12} instead.  3.625 < 5 components,
    so the reference partitions after bucket 5.  Its swap partition
    (util.rs:4-15) leaves [4, 0, 5, 1 | 12].  The left four take the median
    split, and nth_element's quirk (util.rs:39-41) leaves [0, 4, 1, 5]: halves
    {0, 4} and {1, 5}."""
    t = OracleScene(_volume_spheres()).export_bvh()
    assert t["count"].tolist() == [0, 0, 0, 1, 1, 0, 1, 1, 1]
    assert t["offset"].tolist() == [8, 5, 4, 0, 1, 7, 2, 3, 4]
    assert t["axis"].tolist() == [0] * 9
    # x = 0, 4, 1, 5 | 12  ->  prims 2, 1, 4, 3 | 0
    assert t["comp_prim"].tolist() == [2, 1, 4, 3, 0]
    np.testing.assert_array_equal(t["bbox"][1], [-1, -3, -3, 8, 3, 3])
    np.testing.assert_array_equal(t["bbox"][2], [-1, -3, -3, 7, 3, 3])


def test_bvh_sah_leaf_cost_takes_the_median_split():
    """Six spheres of radius 100 at x = 3, 12, 0, 4, 1, 2 (prims 0..5): every box
    spans almost the whole node, so no split is cheaper than a leaf.  Buckets
    are floor(x) = [3, 11, 0, 4, 1, 2].  With equal y/z extents, the volume
    costs are 1 + (n0 (200 + max x left) + n1 (212 - min x right)) / 212:
    6.920, 6.858, 6.816, 6.792 after buckets 0..3, and 6.755 after bucket 4 (the
    minimum).  That is >= 6 components, so the reference takes the "leaf" branch
    (bvh.rs:276-296).  It splits at the median (mid 3) by nth_element rather
    than at bucket 4, which would give {0,1,2,3,4} | {12}.  nth_element over
    x = [3, 12, 0, 4, 1, 2] leaves [0, 2, 3, 1, 4, 12] (its left recursion skips
    an element).  The halves are {0, 2, 3} | {1, 4, 12}, not {0, 1, 2} | {3, 4, 12}.
    Each half of three then splits at its median: {0} | {2, 3} and {1} | {4, 12}."""
    xs = [3, 12, 0, 4, 1, 2]
    spec = _spheres_scene([(x, 0, 0) for x in xs], radius=100.0)
    t = OracleScene(spec).export_bvh()
    assert t["count"].tolist() == [0, 0, 1, 0, 1, 1, 0, 1, 0, 1, 1]
    assert t["offset"].tolist() == [6, 3, 0, 5, 1, 2, 8, 3, 10, 4, 5]
    # x = 0, 2, 3, 1, 4, 12  ->  prims 2, 5, 0, 4, 3, 1
    assert t["comp_prim"].tolist() == [2, 5, 0, 4, 3, 1]


def test_bvh_sah_more_than_255_components_always_split():
    """300 spheres of radius 1e6: 250 at x = 0..249 and 50 at x = 1000..1049
    (prims in that order, reversed so partitioning must move them).  Every
    split costs about 1 + 300 * 0.9999 >= 300 components, so only the
    MAX_COMPONENTS_PER_NODE rule (n > 255, bvh.rs:277) forces an interior
    split at the cheapest bucket.  Buckets: floor(12 x / 1049) is 0, 1 or 2 for
    x < 250 and 11 for x >= 1000.  Up to the shared 2e6 terms, the cost
    numerators n0 max_left + n1 (1049 - min_right) are 211388, 139700 and 64700
    after buckets 0, 1 and 2 (and after 3..10).  So the split falls after
    bucket 2: the 250 near spheres | the 50 far ones.  A leaf-cost test without
    the n > 255 rule would take the median split, 150 | 150."""
    xs = list(range(250)) + list(range(1000, 1050))
    xs = xs[::-1]
    spec = _spheres_scene([(x, 0, 0) for x in xs], radius=1.0e6)
    t = OracleScene(spec).export_bvh()
    left = _subtree_prims(t, 1)
    right = _subtree_prims(t, int(t["offset"][0]))
    assert sorted(np.array(xs)[left].tolist()) == list(range(250))
    assert sorted(np.array(xs)[right].tolist()) == list(range(1000, 1050))
    check_invariants(spec, t)


def _flat_triangles(xs):
    """One small triangle per x in the plane z = 0.  Every box has zero
    volume, so every SAH cost is 1 + 0/0 = NaN."""
    s = scenes.SceneSpec(name="flat")
    verts, tris = [], []
    for k, x in enumerate(xs):
        verts += [(x - 0.3, -0.2, 0.0), (x + 0.3, -0.2, 0.0), (x, 0.4, 0.0)]
        tris.append((3 * k, 3 * k + 1, 3 * k + 2))
    s.add_mesh(np.array(verts, np.float32), tris, scenes.disney())
    return s


def test_bvh_sah_nan_costs():
    """All-NaN costs: the minimum search seeds with cost[0] and compares with a
    strict < (bvh.rs:266-273), so it keeps NaN at bucket 0, and NaN < leaf cost
    is false.  With <= 255 components that is the median split: 100 triangles at
    x = 0..99 split 50 | 50.  Partitioning at bucket 0 would split
    floor(12 x / 99) = 0, x < 8.25: 9 | 91.  With > 255 components the n > 255
    rule partitions at bucket 0: 300 triangles at x = 0..299 split into
    floor(12 x / 299) = 0, x < 24.92: 25 | 275.  A true surface area, non-zero
    for these flat boxes, would give finite costs and split near the middle."""
    for n, n_left in ((100, 50), (300, 25)):
        spec = _flat_triangles(range(n))
        t = OracleScene(spec).export_bvh()
        assert t["count"][0] == 0 and t["axis"][0] == 0
        # the left child's leaves come first in component order; triangle k sits at x = k
        o1 = int(t["offset"][0])
        assert int(sum(t["count"][1:o1])) == n_left
        if n == 300:
            assert sorted(t["comp_index"][:n_left].tolist()) == list(range(n_left))
        check_invariants(spec, t)


def _overflow_spheres():
    # x = 2e14, 1e12, 2e13, 1.99e14, 0 (prims 0..4), radius 1e12
    return _spheres_scene([(2e14, 0, 0), (1e12, 0, 0), (2e13, 0, 0), (1.99e14, 0, 0), (0, 0, 0)],
                          radius=1.0e12)


def test_bvh_sah_nan_first_cost_is_kept():
    """Only cost[0] is NaN; the others are finite.  Boxes are 2e12 wide in y and z,
    so a box longer than 8.5e13 in x has an infinite f32 volume.  The node
    (x extent 2.02e14) is infinite.  After bucket 0 the right side {2e13,
    1.99e14, 2e14} is infinite too, and inf / inf makes cost[0] NaN.  After
    buckets 1..10 both sides are finite, so the cost is 1 + finite / inf = 1.
    The reference seeds the minimum with cost[0] and compares with a strict <
    (bvh.rs:266-273).  So it keeps NaN at bucket 0, NaN < 5 is false, and it
    takes the median split.  nth_element over x = [2e14, 1e12, 2e13, 1.99e14, 0]
    leaves [1e12, 0 | 2e13, 1.99e14, 2e14].  A search seeded with +inf would
    pick cost 1 at bucket 1 and partition {0, 1e12, 2e13} | {1.99e14, 2e14}."""
    t = OracleScene(_overflow_spheres()).export_bvh()
    assert sorted(_subtree_prims(t, 1)) == [1, 4]
    assert sorted(_subtree_prims(t, int(t["offset"][0]))) == [0, 2, 3]


def _sah_cases():
    return {
        "volume": _volume_spheres,
        "leaf_cost": lambda: _spheres_scene([(x, 0, 0) for x in [3, 12, 0, 4, 1, 2]], radius=100.0),
        "over255": lambda: _spheres_scene([(x, 0, 0) for x in (list(range(250)) + list(range(1000, 1050)))[::-1]],
                                          radius=1.0e6),
        "flat100": lambda: _flat_triangles(range(100)),
        "flat300": lambda: _flat_triangles(range(300)),
        "nan_first": _overflow_spheres,
    }


@pytest.mark.parametrize("name", sorted(_sah_cases()))
def test_library_builder_matches_oracle_on_sah_cases(name, host_ctx):
    from gammaray_amd.api import Stage
    spec = _sah_cases()[name]()
    got = Stage(spec, ctx=host_ctx).export_bvh()
    want = OracleScene(spec).export_bvh()
    for k in want:
        assert np.array_equal(got[k], want[k]), k


# --------------------------------------------------------------------------
def _tie_ok(a, b, idx):
    ta, tb = a["dist"][idx], b["dist"][idx]
    return np.all(np.abs(ta - tb) <= 1e-6 * np.maximum(np.abs(ta), np.abs(tb)))


@pytest.mark.parametrize("name", sorted(SCENES))
def test_bvh_matches_brute_force(name):
    spec = SCENES[name]()
    o = OracleScene(spec)
    rays = scenes.rays6(scenes.ray_batch(spec, 20000, seed=11))
    a = o.intersect(rays)
    b = o.intersect_naive(rays)
    assert a["overflow"] == 0
    assert np.array_equal(a["prim"] == NONE, b["prim"] == NONE)
    diff = np.nonzero((a["prim"] != b["prim"]) | (a["comp"] != b["comp"]) | (a["dist"] != b["dist"]))[0]
    # only ties: both hits within 1e-6 relative distance
    assert _tie_ok(a, b, diff)
    assert len(diff) <= 0.001 * len(rays)


def test_closed_form_triangle_and_sphere():
    s = scenes.SceneSpec()
    s.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), [[0, 1, 2]], scenes.disney())
    s.add_sphere((0.0, 0.0, -10.0), 1.0, scenes.disney())
    o = OracleScene(s)
    rays = np.array([
        [0.25, 0.25, 1.0, 0, 0, -1],      # hits the triangle at t = 1
        [0.0, 0.0, -5.0, 0, 0, -1],       # hits sphere front face at t = 4
        [0.0, 0.0, -10.0, 0, 0, 1],       # from the centre: far root t = 1
        [0.25, 0.25, 1.0, 1, 0, 0],       # parallel to the triangle plane: miss
        [2.0, 2.0, 1.0, 0, 0, -1],        # outside the triangle: miss
        [0.5, 0.5, 1.0, 0, 0, -1],        # on edge u = 0 (accepted, u < 0 is the test)
        [0.0, 0.5, 1.0, 0, 0, -1],        # origin on the leaf box plane x = 0 with d.x = 0
    ], np.float32)
    r = o.intersect(rays)
    n = o.intersect_naive(rays)
    assert r["prim"].tolist()[:6] == [0, 1, 1, NONE, NONE, 0]
    assert r["dist"][0] == 1.0 and r["dist"][1] == 4.0 and r["dist"][2] == 1.0
    assert r["dist"][5] == 1.0
    # Q7: (0 - 0) * inf = NaN in the leaf's x slab makes `t_min < max_dist`
    # false, so the tree traversal skips the triangle the brute force hits.
    assert n["prim"][6] == 0 and n["dist"][6] == 1.0
    assert r["prim"][6] == 1 and r["dist"][6] > 10.0


def test_frames_closed_form():
    s = scenes.SceneSpec()
    s.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), [[0, 1, 2]], scenes.disney())
    s.add_sphere((0.0, 0.0, -10.0), 1.0, scenes.disney())
    o = OracleScene(s)
    rays = np.array([[0.25, 0.25, 1.0, 0, 0, -1], [0.0, 0.0, -5.0, 0, 0, -1], [3, -10, 0, -1, 0, 0]], np.float32)
    f = o.intersect(rays, frame=True)["frame"]
    # triangle a=(0,0,0), b=(1,0,0), c=(0,1,0): (a-c) x (b-c) = (0,-1,0) x (1,-1,0) = (0,0,1)
    np.testing.assert_allclose(f[0, 9:12], [0, 0, 1], atol=1e-7)
    # sphere at its +z pole: normal (0,0,1); tangent (-n.z, 0, n.x) normalised = (-1,0,0)
    np.testing.assert_allclose(f[1, 0:3], [0, 0, 1], atol=1e-7)
    np.testing.assert_allclose(f[1, 3:6], [-1, 0, 0], atol=1e-7)
    # frames are orthonormal and right-handed: tangent x binormal = normal
    for row in f[:2]:
        n, t, b = row[0:3], row[3:6], row[6:9]
        np.testing.assert_allclose(np.cross(t, b), n, atol=1e-6)


def test_axis_aligned_rays_match_brute_force():
    # Zero direction components give inv_dir = +-inf and 0 * inf = NaN inside
    # the slab test (src/core/ray.rs:33-36, src/core/bbox.rs:82-112).
    spec = scenes.box_scene()
    o = OracleScene(spec)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.95, 0.95, (3000, 3)).astype(np.float32)
    pts[:, 1] += 1.0
    dirs = np.zeros_like(pts)
    axis = rng.integers(0, 3, len(pts))
    dirs[np.arange(len(pts)), axis] = rng.choice([-1.0, 1.0], len(pts))
    rays = np.concatenate([pts, dirs], 1)
    a = o.intersect(rays)
    b = o.intersect_naive(rays)
    diff = np.nonzero((a["prim"] != b["prim"]) | (a["dist"] != b["dist"]))[0]
    assert _tie_ok(a, b, diff)
    # Origins exactly on a box plane with a zero direction component produce
    # NaN slab distances and the tree may skip hits the brute force finds (Q7);
    # those rays are parity cases for the GPU, not for the brute force.
    pts[::7, 0] = -0.7
    pts[::11, 2] = -0.2
    rays2 = np.concatenate([pts, dirs], 1)
    a2 = o.intersect(rays2)
    b2 = o.intersect_naive(rays2)
    missed = np.nonzero(a2["prim"] != b2["prim"])[0]
    assert len(missed) > 0
    assert np.all((rays2[missed, 0] == -0.7) | (rays2[missed, 2] == np.float32(-0.2)))


def test_visibility_matches_brute_force():
    spec = SCENES["showcase_small"]()
    o = OracleScene(spec)
    segs = scenes.rays6(scenes.segment_batch(spec, 20000, seed=5))
    vis = o.visibility(segs)["visible"]
    start, target = segs[:, :3], segs[:, 3:]
    d = target - start
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    origin = (start + d * np.float32(1e-3)).astype(np.float32)
    tdist = np.linalg.norm(target - origin, axis=1).astype(np.float32)
    b = o.intersect_naive(np.concatenate([origin, d], 1).astype(np.float32))
    occl = (b["prim"] != NONE) & (b["dist"] < tdist - np.float32(1e-3))
    close = np.all(np.abs(target - start) < 1e-3, axis=1)
    want = (~occl & ~close).astype(np.uint8)
    mism = np.nonzero(vis != want)[0]
    # mismatches only where the brute force's and the tree's closest hit tie
    assert len(mism) <= 0.001 * len(segs)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_visibility_early_exit_counts(name):
    """The first-occluder traversal gives the reference's answers.  It does the
    reference's whole traversal on visible segments, since they meet no
    occluder.  On occluded ones it does a prefix of that traversal, and so no
    more box or primitive tests."""
    spec = SCENES[name]()
    o = OracleScene(spec)
    segs = scenes.rays6(scenes.segment_batch(spec, 20000, seed=6))
    full = o.visibility(segs, counts=True)
    early = o.visibility(segs, counts=True, early_exit=True)
    assert np.array_equal(full["visible"], early["visible"])
    vis = full["visible"] == 1
    assert vis.any() and (~vis).any()
    assert np.array_equal(full["counts"][vis], early["counts"][vis])
    assert np.all(early["counts"] <= full["counts"])
    # an occluder ends the traversal early for most occluded segments
    shorter = (early["counts"][~vis].sum(1) < full["counts"][~vis].sum(1)).mean()
    assert shorter > 0.5, shorter


def test_visibility_early_exit_counts_by_hand():
    """Two unit spheres on the x axis at 0 and 3 (a root and two leaves).  A
    segment along +x from x = -5 to x = 10 is blocked by the first sphere.
    The first-occluder traversal tests the root box, then the near child's
    box and its sphere, and stops: 2 box tests and 1 sphere test.  The full
    traversal also pops the far child and tests its box, which starts beyond
    the hit already found (the deferred test of bvh.rs:426), so its sphere is
    not tested: 3 box tests and 1 sphere test."""
    spec = _spheres_scene([(0, 0, 0), (3, 0, 0)], radius=1.0)
    o = OracleScene(spec)
    seg = np.array([[-5, 0, 0, 10, 0, 0]], np.float32)
    early = o.visibility(seg, counts=True, early_exit=True)
    full = o.visibility(seg, counts=True)
    assert early["visible"][0] == 0 and full["visible"][0] == 0
    assert early["counts"][0].tolist() == [2, 0, 1]
    assert full["counts"][0].tolist() == [3, 0, 1]
    # a segment ending before the first sphere is visible: both traverse fully
    # (the sphere is hit beyond the target, so nothing stops the traversal)
    seg2 = np.array([[-5, 0, 0, -2, 0, 0]], np.float32)
    e2 = o.visibility(seg2, counts=True, early_exit=True)
    f2 = o.visibility(seg2, counts=True)
    assert e2["visible"][0] == 1 and f2["visible"][0] == 1
    assert e2["counts"][0].tolist() == f2["counts"][0].tolist()


# --------------------------------------------------------------------------
def _degenerate_scenes():
    out = {}
    out["coincident300"] = _spheres_scene([(0.5, 0.5, 0.5)] * 300)
    out["quirk4"] = _spheres_scene([(1, 0, 0), (3, 0, 0), (9, 0, 0), (2, 0, 0)])
    rng = np.random.default_rng(9)
    pts = rng.uniform(-1, 1, (600, 3)).astype(np.float32)
    pts[:200, 2] = 0.0                      # a planar subset (flat boxes, NaN/inf SAH costs)
    pts[200:260] = pts[200]                 # a coincident cluster inside a larger set
    out["mixed"] = _spheres_scene(pts, radius=0.01)
    out["geometric"] = _spheres_scene([(2.0 ** (-k), 0, 0) for k in range(40)], radius=1e-4)
    out["box"] = scenes.box_scene()
    out["showcase_small"] = scenes.showcase_scene(n_spheres=60, grid=(20, 25))
    out["ico5"] = scenes.icosphere_scene(level=5)
    return out


@pytest.mark.parametrize("name", sorted(_degenerate_scenes()))
def test_library_builder_matches_oracle(name, host_ctx):
    from gammaray_amd.api import Stage
    spec = _degenerate_scenes()[name]
    a = OracleScene(spec).export_bvh()
    b = Stage(spec, ctx=host_ctx).export_bvh()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_abi_exports_every_declared_symbol():
    import os
    from gammaray_amd import _lib
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "gammaray.h")).read()
    declared = set(re.findall(r"\b(gr_[a-z_0-9]+)\s*\(", hdr))
    L = _lib.lib()
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert declared <= set(_lib.SIGNATURES), sorted(declared - set(_lib.SIGNATURES))


def test_host_only_context_rejects_tracing(host_ctx):
    from gammaray_amd.api import Stage
    from gammaray_amd._lib import GrError
    st = Stage(scenes.box_scene(), ctx=host_ctx)
    with pytest.raises(GrError):
        st.intersect_ptr(0, 0, 0)


def test_invalid_mesh_rejected(host_ctx):
    from gammaray_amd.api import Stage
    from gammaray_amd._lib import GrError
    s = scenes.SceneSpec()
    s.add_mesh(np.zeros((3, 3), np.float32), [[0, 1, 5]], scenes.disney())
    with pytest.raises(GrError):
        Stage(s, ctx=host_ctx)


def test_tuning_validation_and_rounding(host_ctx):
    """Sort widths round down to 3 + 3m bits (octant + m bits per axis);
    settings outside the documented ranges are rejected."""
    from gammaray_amd.api import Stage
    from gammaray_amd._lib import GrError
    st = Stage(scenes.box_scene(), ctx=host_ctx)
    assert st.get_tuning() == dict(sort_bits=0, lds_stack=8, waves_per_cu=0, refill=32, xcd_ranges=1,
                                   ray_chunk=128, vis_refill=40)
    for want, got in [(0, 0), (5, 0), (6, 6), (8, 6), (20, 18), (30, 30)]:
        assert st.set_tuning(sort_bits=want)["sort_bits"] == got
    for bad in [dict(sort_bits=31), dict(lds_stack=12), dict(refill=0), dict(refill=65), dict(waves_per_cu=65),
                dict(xcd_ranges=2), dict(ray_chunk=65537), dict(vis_refill=65)]:
        with pytest.raises(GrError):
            st.set_tuning(**bad)
    assert st.get_tuning()["sort_bits"] == 30
    assert st.set_tuning(ray_chunk=1)["ray_chunk"] == 1
    assert st.set_tuning(ray_chunk=65536)["ray_chunk"] == 65536
    with pytest.raises(GrError):
        import torch
        st.autotune(torch.zeros((4, 8)))
    with pytest.raises(GrError):
        st.autotune(torch.zeros((4, 8)), torch.zeros((4, 8)))


def test_batch_rejects_host_only_context(host_ctx):
    from gammaray_amd.api import Stage, render_jobs
    from gammaray_amd._lib import GrError
    st = Stage(scenes.box_scene(), ctx=host_ctx)
    no_film = type("NoFilm", (), {"handle": None})()
    with pytest.raises(GrError):
        render_jobs(host_ctx, [dict(stage=st, film=no_film, spp=1)])


def test_context_outlived_by_its_scene(host_ctx):
    """A scene keeps its context alive: destroying the context's handle first
    and the scene afterwards is safe (the reference's Rust ownership makes
    the opposite order impossible; a C caller or a garbage collector may
    use either)."""
    from gammaray_amd import _lib
    from gammaray_amd.api import Context, Stage
    ctx = Context(_lib.GR_HOST_ONLY)
    st = Stage(scenes.box_scene(), ctx=ctx)
    ctx.close()
    assert st.info()["num_prims"] == len(scenes.box_scene().prims)
    st.close()
