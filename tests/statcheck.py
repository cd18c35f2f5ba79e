"""Statistical checks shared by the GPU parity tests (imported by them; not a
test module).

Per-pixel 3-sigma test of two independent renders: for each pixel, Welch's
statistic z = |mean_a - mean_b| / sqrt(var_a / n_a + var_b / n_b).  Under
equal means each pixel exceeds 3 with probability p0 = 0.0027 (normal
approximation; every test uses at least 64 samples per pixel).  So the number
of exceeding pixels K out of N is Binomial(N, p0).  A render passes when K is
within the binomial's 1 - 1e-4 quantile (about 1.1 % of 1024 pixels), not
within a flat percentage.  A bias confined to a few percent of the pixels
(borders, light edges) then fails.
"""
import numpy as np
from scipy import stats

P0 = 0.0027          # two-sided normal tail beyond 3 sigma
QUANTILE = 1.0 - 1e-4


def pixel_z(a, b):
    """a, b: (samples, pixels) luminance-like sums of two independent renders."""
    na, nb = a.shape[0], b.shape[0]
    se = np.sqrt(a.var(0, ddof=1) / na + b.var(0, ddof=1) / nb)
    return np.abs(a.mean(0) - b.mean(0)) / np.maximum(se, 1e-12)


def exceed_bound(npix, p0=P0, q=QUANTILE):
    """Largest count of pixels beyond 3 sigma accepted among npix."""
    return int(stats.binom.ppf(q, npix, p0))


def assert_three_sigma(a, b, what=""):
    """Per-pixel 3-sigma agreement with the binomial bound, plus the image
    total within 4 sigma.  Returns (pixels beyond 3 sigma, bound)."""
    z = pixel_z(a, b)
    k = int((z > 3.0).sum())
    bound = exceed_bound(z.size)
    print("%s: %d of %d pixels beyond 3 sigma (bound %d)" % (what, k, z.size, bound), flush=True)
    assert k <= bound, (what, k, bound)
    tot = abs(a.mean() - b.mean()) / np.sqrt(a.var() / a.size + b.var() / b.size)
    assert tot < 4.0, (what, tot)
    return k, bound


def assert_three_sigma_blocks(a, b, width, height, block, what=""):
    """assert_three_sigma over block x block pixel tiles instead of pixels:
    a, b (samples, height * width); each tile pools its pixels' samples, which
    suits renders with few samples per pixel.  The pooled variance includes
    the spread between the tile's pixels, so the test is conservative."""
    hb, wb = height // block, width // block

    def tiles(x):
        x = x.reshape(x.shape[0], height, width)[:, :hb * block, :wb * block]
        x = x.reshape(x.shape[0], hb, block, wb, block).transpose(0, 2, 4, 1, 3)
        return x.reshape(x.shape[0] * block * block, hb * wb)

    return assert_three_sigma(tiles(a), tiles(b), what)
