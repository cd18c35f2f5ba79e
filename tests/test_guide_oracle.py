"""The FP64 guide oracle checked against mathematics: finite differences pin
its analytic gradient to its loss, quadrature pins the vMF normalisation,
a chi-square test pins its sampler to its density, and closed forms pin the
input encoding."""
import numpy as np
import pytest

from oracle import guide_oracle as G


def _records(rng, n):
    raw = rng.uniform(0.0, 1.0, (n, 8)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    v = rng.exponential(1.0, n)
    return raw, d, v


def _params(seed, out_scale=8.0):
    p = G.initial_like(seed)
    w0, _ = G.OFFSETS[4]
    p[w0:w0 + 64 * 32] *= out_scale  # a mixture far from uniform
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    params = _params(2)
    raw, d, v = _records(rng, 7)
    g = G.gradient(params, raw, d, v)
    # coordinates in every layer: weights (incl. one of the zero padding rows) and biases
    idx = []
    for (w, b), (nin, nout) in zip(G.OFFSETS, G.LAYERS):
        idx += list(rng.integers(w, w + (min(nin, 200)) * nout, 25)) + list(rng.integers(b, b + nout, 8))
    eps = 1e-6
    worst = 0.0
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (G.loss(p1, raw, d, v) - G.loss(p2, raw, d, v)) / (2 * eps)
        worst = max(worst, abs(fd - g[i]) / max(1e-6, abs(fd) + abs(g[i])))
    assert worst < 1e-4, worst
    assert np.abs(g).max() > 0


def test_padding_rows_get_no_gradient():
    rng = np.random.default_rng(3)
    raw, d, v = _records(rng, 5)
    g = G.gradient(_params(4), raw, d, v)
    w0, _ = G.OFFSETS[0]
    assert np.all(g[w0 + 200 * 64:w0 + 208 * 64] == 0.0)


@pytest.mark.parametrize("kappa", [1e-2, 0.5, 3.0, 30.0, 300.0, 1096.0])
def test_vmf_integrates_to_one(kappa):
    # the density depends on c = mu.w only: integral = 2 pi int_{-1}^{1} pdf(c) dc,
    # with c = 1 - t and t spaced geometrically to resolve the peak at c = 1
    t = np.concatenate([[0.0], np.geomspace(1e-9, 2.0, 200001)])
    c = 1.0 - t
    pdf = G.vmf_pdf(kappa, np.array([0.0, 0.0, 1.0]), np.stack([np.zeros_like(c), np.sqrt(1 - c * c), c], -1))
    integral = 2.0 * np.pi * np.trapezoid(pdf, t)
    assert abs(integral - 1.0) < 1e-6, integral


def test_vmf_small_kappa_is_uniform():
    d = np.array([[0.0, 0.6, 0.8], [0.0, 0.0, -1.0]])
    p = G.vmf_pdf(1e-7, np.array([0.0, 0.0, 1.0]), d)
    np.testing.assert_allclose(p, 1.0 / (4 * np.pi), rtol=1e-6)


def _fibonacci(n):
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (3.0 - np.sqrt(5.0)) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def test_mixture_integrates_to_one():
    rng = np.random.default_rng(5)
    raw, _, _ = _records(rng, 4)
    mix = G.mixture_of(_params(6, out_scale=3.0), raw)
    w, kappa, mu = G.split_mixture(mix)
    assert np.allclose(w.sum(1), 1.0)
    assert kappa.max() < 200  # resolvable by the grid below
    pts = _fibonacci(400_000)
    for i in range(len(mix)):
        q = G.mixture_pdf(np.repeat(w[i:i + 1], len(pts), 0), np.repeat(kappa[i:i + 1], len(pts), 0),
                          np.repeat(mu[i:i + 1], len(pts), 0), pts)
        assert abs(q.mean() * 4 * np.pi - 1.0) < 2e-3


def test_sampler_matches_density():
    """Chi-square of 400k samples in 20 x 40 equal-area cells of (z, phi)."""
    rng = np.random.default_rng(7)
    mix = np.zeros((1, 40))
    lobes = [(0.5, 4.0, (0, 0, 1)), (0.3, 20.0, (1, 0, 0)), (0.2, 0.5, (0, -0.6, -0.8))]
    for i in range(8):
        wt, k, m = lobes[i] if i < 3 else (0.0, 1.0, (0, 0, 1))
        mix[0, 5 * i:5 * i + 5] = [wt, k, *m]
    n = 400_000
    u = rng.random((n, 3))
    s = G.sample_mixture(np.repeat(mix, n, 0), u)
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0)
    nz, nphi = 20, 40
    zi = np.minimum(((s[:, 2] + 1) / 2 * nz).astype(int), nz - 1)
    pi_ = np.minimum(((np.arctan2(s[:, 1], s[:, 0]) + np.pi) / (2 * np.pi) * nphi).astype(int), nphi - 1)
    counts = np.bincount(zi * nphi + pi_, minlength=nz * nphi)
    # expected cell mass by fine quadrature of the density
    fz, fp = 40, 40
    z = -1 + (np.arange(nz * fz) + 0.5) * (2.0 / (nz * fz))
    ph = -np.pi + (np.arange(nphi * fp) + 0.5) * (2 * np.pi / (nphi * fp))
    Z, P = np.meshgrid(z, ph, indexing="ij")
    R = np.sqrt(1 - Z * Z)
    pts = np.stack([R * np.cos(P), R * np.sin(P), Z], -1).reshape(-1, 3)
    w, kappa, mu = G.split_mixture(mix)
    q = G.mixture_pdf(np.repeat(w, len(pts), 0), np.repeat(kappa, len(pts), 0), np.repeat(mu, len(pts), 0), pts)
    cell = (2.0 / (nz * fz)) * (2 * np.pi / (nphi * fp))
    mass = (q * cell).reshape(nz, fz, nphi, fp).sum((1, 3)).ravel()
    assert abs(mass.sum() - 1.0) < 1e-3
    exp = mass * n
    keep = exp > 5
    chi2 = ((counts[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    dof = keep.sum() - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)


def test_encoding_closed_forms():
    raw = np.zeros((3, 8), np.float32)
    raw[1, 0] = 0.5
    raw[2, 3] = 0.25
    f = G.encode(raw)
    assert np.allclose(f[0, 0:192:2], 0.0) and np.allclose(f[0, 1:192:2], 1.0)
    # x = 0.5: sin(pi/2) = 1, cos = 0 at k = 0; at k >= 1 the argument is a whole turn multiple
    assert np.isclose(f[1, 0], 1.0) and np.isclose(f[1, 1], 0.0, atol=1e-12)
    assert np.isclose(f[1, 2], 0.0, atol=1e-12) and np.isclose(f[1, 3], -1.0)
    # x = 0.25 in dimension 3: sin(pi/4), sin(pi/2), sin(pi) at k = 0, 1, 2
    base = 24 * 3
    np.testing.assert_allclose(f[2, [base, base + 2, base + 4]], [np.sqrt(0.5), 1.0, 0.0], atol=1e-12)
    assert f[2, 192 + 3] == np.float32(0.25) and np.all(f[:, 200:] == 0.0)
