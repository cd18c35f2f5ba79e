"""FP64 CPU restatement of the learned path guide -- TEST INFRASTRUCTURE ONLY.

The guide has no counterpart in sdao/gammaray (nothing there learns or
guides), so this module is the specification the gfx950 kernels are checked
against.  Only tests/, the smoke check and bench.py's CPU leg may use it.

Network: 8 raw inputs -> sin/cos(2^k pi x), k = 0..11 (192) + the 8 raw values,
zero-padded to 208 -> 64 -> 64 -> 64 -> 64 (ReLU) -> 32 outputs = 8 lobes x
(weight logit, log kappa, theta logit, phi logit).  Lobe i is a von
Mises-Fisher density with weight softmax(logits)_i, kappa = exp(clip(log
kappa, -4, 7)), mean direction (sin t cos p, sin t sin p, cos t) with
t = pi sigmoid, p = 2 pi sigmoid.

Training loss per record (raw, direction w, value v): -(v / mean v) log q(w),
with v / mean v clipped to [0, 20]: a Monte Carlo estimate of the KL
divergence (up to a constant) from the distribution proportional to
f*cos*Li -- whose samples/pdf the renderer reports as v -- to the mixture q.

Pinned by: central finite differences of `loss` against `gradient`
(tests/test_guide_oracle.py), quadrature of every vMF lobe and of whole
mixtures to 1, a chi-square test of `sample_mixture` against `mixture_pdf`,
and limits (kappa -> 0 is uniform, 1/(4 pi)).
"""
import numpy as np

RAW, FREQS, ENCODED, IN, WIDTH, OUT, LOBES = 8, 12, 192, 208, 64, 32, 8
LAYERS = [(IN, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, OUT)]
LOG_KAPPA_LO, LOG_KAPPA_HI = -4.0, 7.0
VALUE_CLIP = 20.0


def _offsets():
    off, o = [], 0
    for nin, nout in LAYERS:
        off.append((o, o + nin * nout))
        o += nin * nout + nout
    return off, o


OFFSETS, NPARAMS = _offsets()
assert NPARAMS == 27936


def unpack(params):
    params = np.asarray(params, np.float64)
    out = []
    for (w, b), (nin, nout) in zip(OFFSETS, LAYERS):
        out.append((params[w:w + nin * nout].reshape(nin, nout), params[b:b + nout]))
    return out


def pack(layers):
    p = np.zeros(NPARAMS)
    for (w, b), (W, B) in zip(OFFSETS, layers):
        p[w:w + W.size] = W.ravel()
        p[b:b + B.size] = B
    return p


def encode(raw):
    """(n, 8) raw inputs (float32 values) -> (n, 208) features."""
    raw = np.asarray(raw, np.float32).astype(np.float64)
    n = len(raw)
    f = np.zeros((n, IN))
    for d in range(RAW):
        for k in range(FREQS):
            a = raw[:, d] * 2.0 ** k
            r = a - 2.0 * np.floor(a * 0.5)
            f[:, 24 * d + 2 * k] = np.sin(np.pi * r)
            f[:, 24 * d + 2 * k + 1] = np.cos(np.pi * r)
    f[:, ENCODED:ENCODED + RAW] = raw
    return f


def _h(x, f16):
    return x.astype(np.float16).astype(np.float64) if f16 else x


def forward(params, raw, f16=False):
    """Outputs (n, 32) and the activations [features, h1..h4].  f16=True rounds
    weights, features and hidden activations to fp16 where the kernels do."""
    layers = unpack(params)
    a = _h(encode(raw), f16)
    acts = [a]
    for li, (W, B) in enumerate(layers):
        z = a @ _h(W, f16) + B
        if li < 4:
            a = _h(np.maximum(z, 0.0), f16)
            acts.append(a)
        else:
            a = z
    return a, acts


def head(out):
    """Mixture (weights (n,8), kappa (n,8), mu (n,8,3)) and intermediates."""
    o = out.reshape(-1, LOBES, 4)
    logit, lk, ta, pa = o[..., 0], o[..., 1], o[..., 2], o[..., 3]
    m = logit.max(1, keepdims=True)
    e = np.exp(logit - m)
    w = e / e.sum(1, keepdims=True)
    kappa = np.exp(np.clip(lk, LOG_KAPPA_LO, LOG_KAPPA_HI))
    sa, sb = 1.0 / (1.0 + np.exp(-ta)), 1.0 / (1.0 + np.exp(-pa))
    th, ph = np.pi * sa, 2.0 * np.pi * sb
    mu = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    return dict(w=w, kappa=kappa, mu=mu, lk=lk, sa=sa, sb=sb, th=th, ph=ph)


def vmf_pdf(kappa, mu, d):
    """kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.d - 1))."""
    kappa = np.asarray(kappa, np.float64)
    return kappa / (2.0 * np.pi * -np.expm1(-2.0 * kappa)) * np.exp(kappa * ((mu * d).sum(-1) - 1.0))


def mixture_pdf(w, kappa, mu, d):
    """d: (n, 3) -> (n,)"""
    return (w * vmf_pdf(kappa, mu, np.asarray(d, np.float64)[:, None, :])).sum(1)


def mixture_of(params, raw, f16=False):
    """(n, 40) in the kernels' layout: lobe i -> (w, kappa, mu.x, mu.y, mu.z)."""
    out, _ = forward(params, raw, f16)
    h = head(out)
    mix = np.concatenate([h["w"][..., None], h["kappa"][..., None], h["mu"]], -1)
    return mix.reshape(len(raw), 5 * LOBES)


def split_mixture(mix):
    m = np.asarray(mix, np.float64).reshape(-1, LOBES, 5)
    return m[..., 0], m[..., 1], m[..., 2:5]


def sample_mixture(mix, u):
    """Directions from mixtures (n, 40) with uniforms (n, 3): lobe by inverse
    CDF of the weights, cosine to the mean by inversion, uniform azimuth."""
    w, kappa, mu = split_mixture(mix)
    u = np.asarray(u, np.float64)
    n = len(u)
    cdf = np.cumsum(w, 1)
    sel = np.minimum((u[:, :1] >= cdf).sum(1), LOBES - 1)
    k = kappa[np.arange(n), sel]
    m = mu[np.arange(n), sel]
    e = np.exp(-2.0 * k)
    c = np.clip(1.0 + np.log(e + u[:, 1] * (1.0 - e)) / k, -1.0, 1.0)
    s = np.sqrt(np.maximum(0.0, 1.0 - c * c))
    ph = 2.0 * np.pi * u[:, 2]
    sg = np.where(m[:, 2] >= 0.0, 1.0, -1.0)
    aa = -1.0 / (sg + m[:, 2])
    bb = m[:, 0] * m[:, 1] * aa
    t = np.stack([1.0 + sg * m[:, 0] ** 2 * aa, sg * bb, -sg * m[:, 0]], 1)
    b = np.stack([bb, sg + m[:, 1] ** 2 * aa, -m[:, 1]], 1)
    return m * c[:, None] + (t * np.cos(ph)[:, None] + b * np.sin(ph)[:, None]) * s[:, None]


def _values(values):
    v = np.maximum(np.asarray(values, np.float64), 0.0)
    mean = v.mean()
    return np.clip(v / mean, 0.0, VALUE_CLIP) if mean > 0 else np.zeros_like(v)


def loss(params, raw, dirs, values):
    out, _ = forward(params, raw)
    h = head(out)
    q = mixture_pdf(h["w"], h["kappa"], h["mu"], dirs)
    return float(np.mean(-_values(values) * np.log(q)))


def gradient(params, raw, dirs, values):
    """Analytic gradient of `loss` (back-propagation in fp64)."""
    layers = unpack(params)
    out, acts = forward(params, raw)
    h = head(out)
    d = np.asarray(dirs, np.float64)
    n = len(d)
    vn = _values(values)
    pdf = vmf_pdf(h["kappa"], h["mu"], d[:, None, :])
    pw = h["w"] * pdf
    q = pw.sum(1, keepdims=True)
    r = pw / q
    s = (-vn / n)[:, None]  # d(mean loss)/d(log q) per record
    k = h["kappa"]
    md = (h["mu"] * d[:, None, :]).sum(-1) - 1.0
    inside = (h["lk"] > LOG_KAPPA_LO) & (h["lk"] < LOG_KAPPA_HI)
    dlogit = s * (r - h["w"])
    dlk = np.where(inside, s * r * (1.0 - 2.0 * k / np.expm1(2.0 * k) + k * md), 0.0)
    th, ph = h["th"], h["ph"]
    dmu_t = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)], -1)
    dmu_p = np.stack([-np.sin(th) * np.sin(ph), np.sin(th) * np.cos(ph), np.zeros_like(th)], -1)
    dta = s * r * k * (d[:, None, :] * dmu_t).sum(-1) * np.pi * h["sa"] * (1.0 - h["sa"])
    dpa = s * r * k * (d[:, None, :] * dmu_p).sum(-1) * 2.0 * np.pi * h["sb"] * (1.0 - h["sb"])
    dz = np.stack([dlogit, dlk, dta, dpa], -1).reshape(n, OUT)
    grads = [None] * 5
    for li in range(4, -1, -1):
        W, _ = layers[li]
        a_in = acts[li]
        grads[li] = (a_in.T @ dz, dz.sum(0))
        if li > 0:
            dz = (dz @ W.T) * (acts[li] > 0.0)
    return pack(grads)


def initial_like(seed):
    """Random parameters of the kernels' initial scale (for tests)."""
    rng = np.random.default_rng(seed)
    layers = []
    for li, (nin, nout) in enumerate(LAYERS):
        fan = ENCODED + RAW if li == 0 else nin
        bound = (0.1 if li == 4 else 1.0) * np.sqrt(6.0 / fan)
        W = rng.uniform(-bound, bound, (nin, nout))
        if li == 0:
            W[ENCODED + RAW:] = 0.0
        layers.append((W, rng.uniform(-0.05, 0.05, nout)))
    return pack(layers)
