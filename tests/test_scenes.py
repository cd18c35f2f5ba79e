"""CPU tests of the synthetic workload generators the bench tunes on, and of
the guide settings' host-side checks."""
import numpy as np

from gammaray_amd import scenes


def _terrain():
    specs, _ = scenes.batch_workload(seed=4)
    return specs[1]


def test_bounce_rays_leave_the_surface_cosine_distributed():
    spec = _terrain()
    n = 200000
    r = scenes.bounce_batch(spec, n, seed=9)
    rng = np.random.default_rng(9)
    pts, nrm = scenes._surface_points(spec, rng, n)
    d = r[:, 4:7].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    cos = (d * nrm).sum(1)
    assert cos.min() > 0.0
    # cosine-distributed about the normal: cos theta has density 2 c on [0, 1],
    # mean 2/3 and P(cos < 1/2) = 1/4
    assert abs(cos.mean() - 2.0 / 3.0) < 3e-3
    assert abs((cos < 0.5).mean() - 0.25) < 3e-3
    # origins 1e-4 above surface points inside the terrain's footprint
    np.testing.assert_allclose(r[:, 0:3], pts + 1e-4 * nrm, atol=1e-6)
    lo, hi = spec.bbox()
    assert np.all(r[:, 0:3] >= lo - 1e-3) and np.all(r[:, 0:3] <= hi + 1e-3)


def test_light_segments_end_on_the_light():
    spec = _terrain()
    g = scenes.light_segment_batch(spec, 50000, seed=10)
    light = [p for p in spec.prims if p.kind == "mesh" and p.material[0] == 1][0]
    lo, hi = light.vertices.min(0), light.vertices.max(0)
    t = g[:, 4:7]
    assert np.all(t >= lo - 1e-5) and np.all(t <= hi + 1e-5)
    # the quad light is horizontal: every target at its height, spread over it
    np.testing.assert_allclose(t[:, 1], lo[1], atol=1e-5)
    assert t[:, 0].std() > 0.25 * (hi[0] - lo[0]) and t[:, 2].std() > 0.25 * (hi[2] - lo[2])
    assert np.all(g[:, 1] < lo[1])  # starts below the light


def test_guide_settings_are_checked_before_the_device():
    """Unknown lobe sets and lobe choices are refused by the Python layer, and
    a host-only context refuses a guide (no device to hold it)."""
    import pytest
    from gammaray_amd import _lib
    from gammaray_amd.api import Context, Guide
    ctx = Context(_lib.GR_HOST_ONLY)
    with pytest.raises(ValueError):
        Guide(ctx, lobes="transmission")
    with pytest.raises(ValueError):
        Guide(ctx, lobe_select="best")
    with pytest.raises(_lib.GrError):
        Guide(ctx, lobe_select="learned", select_mix=0.4, rr_threshold=0.1)
