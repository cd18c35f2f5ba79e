"""The traversal kernels' interval form of the box test (slab_plain in
gammaray_amd/csrc/gr_traverse.hip) against the reference's staged form
(BBox::intersect, src/core/bbox.rs:78-113, restated by the oracle).

The kernels use the interval form only for a plain ray (origin and 1/d finite,
1/d nonzero on every axis) against a scene whose boxes are all finite with
lo <= hi.  Under those conditions it must agree with the staged form on the
hit decision and on the entry distance for every max distance: checked here on
random, boundary and overflowing inputs, in float32 with one rounding per
operation as on the GPU.
"""
import numpy as np
import pytest

from oracle import oracle


def _pad():
    eps = np.float32(np.finfo(np.float32).eps)
    n = np.float32(3.0)
    g = (n * eps) / (np.float32(1.0) - n * eps)
    return np.float32(1.0) + np.float32(2.0) * g


def _interval_form(boxes, rays, z_raises_entry=False):
    """Returns (pass, entry) of the interval test; z_raises_entry=True is a
    deliberately wrong variant used to show the cases can tell them apart."""
    lo, hi = boxes[:, :3], boxes[:, 3:]
    o, d = rays[:, :3], rays[:, 3:]
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        inv = np.float32(1.0) / d
        neg = d < 0
        near = np.where(neg, hi, lo)
        far = np.where(neg, lo, hi)
        t_near = (near - o) * inv
        t_far = (far - o) * inv
        entry = np.maximum(t_near[:, 0], t_near[:, 1])
        if z_raises_entry:
            entry = np.maximum(entry, t_near[:, 2])
        exit_ = np.minimum(np.minimum(t_far[:, 0], t_far[:, 1]), t_far[:, 2]) * _pad()
        ok = (np.maximum(entry, t_near[:, 2]) <= exit_) & (exit_ > 0)
    return ok, entry


def _plain(boxes, rays):
    with np.errstate(divide="ignore"):
        inv = np.float32(1.0) / rays[:, 3:]
    ok_ray = np.isfinite(rays[:, :3]).all(1) & np.isfinite(inv).all(1) & (inv != 0).all(1)
    ok_box = np.isfinite(boxes).all(1) & (boxes[:, :3] <= boxes[:, 3:]).all(1)
    return ok_ray & ok_box


def _cases(seed, n):
    rng = np.random.default_rng(seed)
    f = np.float32
    lo = rng.uniform(-10, 10, (n, 3)).astype(f)
    ext = (rng.uniform(0, 5, (n, 3)) * (rng.random((n, 3)) > 0.1)).astype(f)  # 10 % flat axes
    hi = (lo + ext).astype(f)
    o = rng.uniform(-20, 20, (n, 3)).astype(f)
    # origins on a face, an edge or a corner of their box
    for k in range(3):
        on = rng.random(n) < 0.15
        o[on, k] = np.where(rng.random(on.sum()) < 0.5, lo[on, k], hi[on, k])
    d = rng.normal(size=(n, 3)).astype(f)
    # tiny but normal direction components (huge finite 1/d)
    tiny = rng.random((n, 3)) < 0.05
    d[tiny] = np.copysign(f(1e-30), d[tiny])
    # rays aimed at the box centre
    aim = rng.random(n) < 0.3
    d[aim] = ((lo[aim] + hi[aim]) * f(0.5) - o[aim]).astype(f)
    boxes = np.concatenate([lo, hi], 1)
    rays = np.concatenate([o, d], 1)
    # overflowing magnitudes: coordinates near FLT_MAX
    big = rng.random(n) < 0.05
    s = f(1e37)
    boxes[big] = (boxes[big] * s).astype(f)
    rays[big, :3] = (rays[big, :3] * f(-1.5e37)).astype(f)
    return boxes.astype(f), rays.astype(f)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_interval_form_matches_staged_box_test(seed):
    boxes, rays = _cases(seed, 200_000)
    keep = _plain(boxes, rays)
    assert keep.mean() > 0.9
    boxes, rays = boxes[keep], rays[keep]
    ok, entry = _interval_form(boxes, rays)
    assert ok.any() and (~ok).any()
    inf = np.full(len(boxes), np.inf, np.float32)
    # hit decision with no distance bound, then the entry distance itself: a
    # bound equal to it rejects, the next float above it accepts
    assert np.array_equal(oracle.bbox_intersect(boxes, rays, inf), ok & (entry < np.inf))
    fin = ok & np.isfinite(entry)
    at = np.where(fin, entry, np.float32(0)).astype(np.float32)
    above = np.nextafter(at, np.float32(np.inf)).astype(np.float32)
    assert not oracle.bbox_intersect(boxes[fin], rays[fin], at[fin]).any()
    assert oracle.bbox_intersect(boxes[fin], rays[fin], above[fin]).all()


def test_cases_tell_the_z_slab_rule_apart():
    """The staged test never lets the z slab raise the entry distance; an
    interval form that did would be caught by the cases above."""
    boxes, rays = _cases(4, 200_000)
    keep = _plain(boxes, rays)
    boxes, rays = boxes[keep], rays[keep]
    ok, entry = _interval_form(boxes, rays)
    ok_z, entry_z = _interval_form(boxes, rays, z_raises_entry=True)
    differs = ok & (entry != entry_z)
    assert differs.any()
    # the oracle sides with the rule that leaves the entry at max(tx, ty)
    at = entry[differs]
    above = np.nextafter(at, np.float32(np.inf)).astype(np.float32)
    assert oracle.bbox_intersect(boxes[differs], rays[differs], above).all()
