"""GPU parity of the learned path guide against its FP64 oracle, and of the
guided integrator against plain PT+NEE.

  * inference (fp16 MFMA, fp32 accumulation) matches the oracle run with the
    same fp16 roundings, to the fp32-accumulation-order level;
  * the training gradient (forward, head, back-propagation, split-K weight
    gradient) matches the oracle's FP64 gradient per layer;
  * Adam's first step moves each weight by -lr sign(gradient);
  * the mixture sampler draws from the density the pdf hook reports;
  * training fits a direction distribution that depends on the input;
  * guided rendering has the plain estimator's mean (per-pixel 3 sigma) on
    pdf-consistent materials, and lower variance where light arrives
    indirectly.
"""
import numpy as np
import pytest

from gammaray_amd import scenes
from gammaray_amd.api import Film, Guide, Stage, mixture_pdf, mixture_sample, trace
from oracle import guide_oracle as G
from statcheck import assert_three_sigma, assert_three_sigma_blocks

pytestmark = pytest.mark.gpu


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _params(seed, out_scale):
    p = G.initial_like(seed)
    w0, _ = G.OFFSETS[4]
    p[w0:w0 + 64 * 32] *= out_scale
    return p.astype(np.float32)


def _records(rng, n, dirs=None):
    raw = rng.uniform(0.0, 1.0, (n, 8)).astype(np.float32)
    if dirs is None:
        dirs = rng.normal(size=(n, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rec = np.zeros((n, 12), np.float32)
    rec[:, :8] = raw
    rec[:, 8:11] = dirs
    rec[:, 11] = rng.exponential(1.0, n)
    return rec


def _network_order(mix):
    """The inference kernel writes a lane's two lobes as one run: mixture slot
    2i holds the network's lobe i and slot 2i + 1 its lobe i + 4 (the order
    of the lobes within a mixture does not matter to sampling or density).
    Back to the network's order, to compare with the oracle lobe by lobe."""
    m = np.asarray(mix).reshape(-1, 8, 5)
    return m[:, [0, 2, 4, 6, 1, 3, 5, 7], :].reshape(-1, 40)


def test_inference_matches_oracle(gpu_ctx):
    print("case guide inference", flush=True)
    rng = np.random.default_rng(1)
    g = Guide(gpu_ctx)
    p = _params(11, 4.0)
    g.set_params(p)
    assert np.array_equal(g.get_params(), p)
    raw = rng.uniform(0.0, 1.0, (4099, 8)).astype(np.float32)
    mix = g.infer(_cuda(raw)).cpu().numpy()
    ref = G.mixture_of(p.astype(np.float64), raw, f16=True)
    w, k, mu = G.split_mixture(_network_order(mix))
    wr, kr, mr = G.split_mixture(ref)
    np.testing.assert_allclose(w.sum(1), 1.0, atol=1e-5)
    np.testing.assert_allclose(w, wr, atol=5e-3)
    np.testing.assert_allclose(k, kr, rtol=2e-2)
    np.testing.assert_allclose(mu, mr, atol=1e-2)
    d = rng.normal(size=(len(raw), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    qg = mixture_pdf(gpu_ctx, _cuda(mix), _cuda(d)).cpu().numpy()
    np.testing.assert_allclose(qg, G.mixture_pdf(w, k, mu, d), rtol=1e-4, atol=1e-7)
    np.testing.assert_allclose(qg, G.mixture_pdf(wr, kr, mr, d), rtol=6e-2, atol=1e-4)


@pytest.mark.parametrize("n", [1, 15, 16, 17, 200_003])
def test_inference_ragged_counts(n, gpu_ctx):
    """Counts that are not whole 16-query tiles: every query is written, nothing past n is."""
    print("case guide inference ragged", n, flush=True)
    import torch
    rng = np.random.default_rng(n)
    g = Guide(gpu_ctx)
    p = _params(14, 4.0)
    g.set_params(p)
    raw = rng.uniform(0.0, 1.0, (n, 8)).astype(np.float32)
    out = torch.full((n + 16, 40), -7.0, dtype=torch.float32, device="cuda")
    from gammaray_amd._lib import check, lib
    import ctypes
    check(lib().gr_guide_infer(g._h, ctypes.c_void_p(_cuda(raw).data_ptr()), n, ctypes.c_void_p(out.data_ptr())))
    got = out.cpu().numpy()
    assert np.all(got[n:] == -7.0)
    sel = rng.choice(n, min(n, 512), replace=False)
    ref = G.mixture_of(p.astype(np.float64), raw[sel], f16=True)
    w, k, mu = G.split_mixture(_network_order(got[sel]))
    wr, kr, mr = G.split_mixture(ref)
    np.testing.assert_allclose(w, wr, atol=5e-3)
    np.testing.assert_allclose(k, kr, rtol=2e-2)
    np.testing.assert_allclose(mu, mr, atol=1e-2)


def test_inference_rejects_misaligned_buffers(gpu_ctx):
    import torch
    from gammaray_amd._lib import GrError
    g = Guide(gpu_ctx)
    flat = torch.zeros(8 * 32 + 1, dtype=torch.float32, device="cuda")
    with pytest.raises(GrError):
        g.infer(flat[1:].view(32, 8))


def test_gradient_matches_oracle(gpu_ctx):
    print("case guide gradient", flush=True)
    rng = np.random.default_rng(2)
    g = Guide(gpu_ctx)
    p = _params(12, 4.0)
    g.set_params(p)
    rec = _records(rng, 2051)
    got = g.gradient(_cuda(rec)).cpu().numpy().astype(np.float64)
    ref = G.gradient(p.astype(np.float64), rec[:, :8], rec[:, 8:11].astype(np.float64), rec[:, 11])
    for li, ((w, b), (nin, nout)) in enumerate(zip(G.OFFSETS, G.LAYERS)):
        for name, lo, hi in (("W", w, w + nin * nout), ("b", b, b + nout)):
            a, r = got[lo:hi], ref[lo:hi]
            rel = np.linalg.norm(a - r) / np.linalg.norm(r)
            cos = a @ r / (np.linalg.norm(a) * np.linalg.norm(r))
            assert rel < 5e-2 and cos > 0.998, (li, name, rel, cos)


def test_adam_first_step(gpu_ctx):
    print("case guide adam", flush=True)
    rng = np.random.default_rng(3)
    lr = 1e-3
    g = Guide(gpu_ctx, learning_rate=lr, batch=65536)
    p = _params(13, 4.0)
    g.set_params(p)
    rec = _records(rng, 2048)
    g.train(_cuda(rec), 1)
    assert g.steps() == 1
    step = g.get_params().astype(np.float64) - p
    assert np.abs(step).max() <= lr * 1.01
    ref = G.gradient(p.astype(np.float64), rec[:, :8], rec[:, 8:11].astype(np.float64), rec[:, 11])
    big = np.abs(ref) > np.quantile(np.abs(ref[ref != 0]), 0.5)
    agree = np.mean(np.sign(step[big]) == -np.sign(ref[big]))
    assert agree > 0.95, agree
    assert np.mean(np.abs(np.abs(step[big]) - lr) < 0.05 * lr) > 0.9


def test_mixture_sampler_on_gpu(gpu_ctx):
    print("case guide sampler", flush=True)
    rng = np.random.default_rng(4)
    mix = np.zeros((1, 40), np.float32)
    lobes = [(0.5, 4.0, (0, 0, 1)), (0.3, 20.0, (1, 0, 0)), (0.2, 0.5, (0, -0.6, -0.8))]
    for i in range(8):  # unused lobes: zero weight, a kappa the network can produce
        wt, k, m = lobes[i] if i < 3 else (0.0, 1.0, (0, 0, 1))
        mix[0, 5 * i:5 * i + 5] = [wt, k, *m]
    n = 400_000
    M = np.repeat(mix, n, 0)
    u = rng.random((n, 3)).astype(np.float32)
    out = mixture_sample(gpu_ctx, _cuda(M), _cuda(u)).cpu().numpy()
    d = out[:, :3].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    w, k, mu = G.split_mixture(M)
    q = G.mixture_pdf(w, k, mu, d)
    assert np.isfinite(out).all()
    np.testing.assert_allclose(out[:, 3], q, rtol=1e-4)
    # the same uniforms through the oracle's sampler give the same directions
    ref = G.sample_mixture(M, u.astype(np.float64))
    assert np.mean(np.linalg.norm(ref - d, axis=1) < 1e-4) > 0.999
    # importance-sampling identity: E[1 / q] = 4 pi
    est = np.mean(1.0 / q)
    assert abs(est / (4 * np.pi) - 1.0) < 0.02, est


def test_training_fits_an_input_dependent_distribution(gpu_ctx):
    """Directions concentrate around +z when raw[0] < 0.5 and around +x
    otherwise; training must raise the held-out log-likelihood far above the
    uniform mixture's log(1 / 4 pi)."""
    print("case guide training", flush=True)
    import torch
    rng = np.random.default_rng(5)
    n = 1 << 16

    def draw(m):
        raw = rng.uniform(0.0, 1.0, (m, 8)).astype(np.float32)
        tgt = np.zeros((m, 40))
        up = raw[:, 0] < 0.5
        tgt[:, 0] = 1.0
        tgt[:, 1] = 30.0
        tgt[:, 2:5] = np.where(up[:, None], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0])
        d = G.sample_mixture(tgt, rng.random((m, 3)))
        rec = np.zeros((m, 12), np.float32)
        rec[:, :8], rec[:, 8:11], rec[:, 11] = raw, d, 1.0
        return rec

    train, held = draw(n), draw(8192)
    g = Guide(gpu_ctx, learning_rate=3e-3, batch=4096, seed=7)

    def loglik():
        mix = g.infer(_cuda(held[:, :8]))
        q = mixture_pdf(gpu_ctx, mix, _cuda(held[:, 8:11])).cpu().numpy()
        return float(np.mean(np.log(np.maximum(q, 1e-30))))

    before = loglik()
    g.train(_cuda(train), 600)
    torch.cuda.synchronize()
    after = loglik()
    assert g.steps() == 600
    assert after > before + 2.0 and after > np.log(1 / (4 * np.pi)) + 2.5, (before, after)


def _sample_radiance(stage, film, integrator, spp, **kw):
    import torch
    W, H = film.width, film.height
    buf = torch.zeros((spp * W * H, 4), dtype=torch.float32, device="cuda")
    trace(stage, film, integrator, spp=spp, sample_radiance_ptr=buf.data_ptr(), **kw)
    stage.ctx.synchronize()
    return buf.cpu().numpy()[:, :3].reshape(spp, H * W, 3)


@pytest.mark.parametrize("variant,lobes,select", [("consistent", "consistent", "uniform"),
                                                  ("clearcoat", "consistent", "uniform"),
                                                  ("transmission", "consistent", "uniform"),
                                                  ("clearcoat", "all", "uniform"),
                                                  ("transmission", "all", "uniform"),
                                                  ("transmission", "consistent", "albedo"),
                                                  ("clearcoat", "all", "albedo"),
                                                  ("transmission", "all", "albedo"),
                                                  ("transmission", "consistent", "learned"),
                                                  ("clearcoat", "all", "learned"),
                                                  ("transmission", "consistent", "learned+roulette"),
                                                  ("transmission", "diffuse", "learned+roulette"),
                                                  ("consistent", "consistent", "uniform+roulette")])
def test_guided_render_has_the_plain_mean(variant, lobes, select, gpu_ctx):
    """Guided and plain PT+NEE per-pixel means agree within 3 sigma (different
    seeds), with the guide sampled after warm-up -- also where clearcoat and
    transmission lobes report pdfs that are not their sampling densities, and
    with those lobes' picks guided too (lobes="all": weighted with the
    densities their samplers realise and the radiance sample_world reports
    for each pick, the extra lobe values of transmitted picks included), with
    the lobe chosen by albedo or by learned statistics rather than uniformly,
    and with the guided integrator's Russian roulette (rr_threshold 0.3, so
    it fires often on these boxes)."""
    print("case guided unbiased", variant, lobes, select, flush=True)
    spec = {"consistent": lambda: scenes.box_scene(consistent=True),
            "clearcoat": lambda: scenes.box_scene(),
            "transmission": lambda: scenes.box_scene(seed=2, full=True)}[variant]()
    W, H, spp = 32, 32, 256
    stage = Stage(spec, ctx=gpu_ctx)
    choice, _, rr = select.partition("+")
    g = Guide(gpu_ctx, batch=4096, steps_per_pass=4, train_paths=4096, warmup_steps=8, alpha=0.5, lobes=lobes,
              lobe_select=choice, rr_threshold=0.3 if rr else 0.0)
    film = Film(gpu_ctx, W, H)
    a = _sample_radiance(stage, film, "guided", spp, max_depth=6, seed=31, batch=8, guide=g)
    assert g.steps() == 4 * spp // 8
    film2 = Film(gpu_ctx, W, H)
    b = _sample_radiance(stage, film2, "pt_nee", spp, max_depth=6, seed=32)
    assert np.isfinite(a).all()
    # iterations traced with the guide sampled
    assert_three_sigma(a[32:].sum(2), b[32:].sum(2), "guided %s %s %s" % (variant, lobes, select))


def test_guide_rejects_unknown_lobe_settings(gpu_ctx):
    from gammaray_amd import _lib
    from gammaray_amd.api import lib
    import ctypes
    p = _lib.GuideParams(1, 5e-3, 0.5, 4096, 4, 4096, 0, 8, 3)
    h = ctypes.c_void_p()
    assert lib().gr_guide_create(gpu_ctx.handle, ctypes.byref(p), ctypes.byref(h)) != 0
    assert not h.value
    with pytest.raises(ValueError):
        Guide(gpu_ctx, lobes="clearcoat")
    p = _lib.GuideParams(1, 5e-3, 0.5, 4096, 4, 4096, 0, 8, 0, 3, 0.25)
    assert lib().gr_guide_create(gpu_ctx.handle, ctypes.byref(p), ctypes.byref(h)) != 0
    p = _lib.GuideParams(1, 5e-3, 0.5, 4096, 4, 4096, 0, 8, 0, 1, 0.0)
    assert lib().gr_guide_create(gpu_ctx.handle, ctypes.byref(p), ctypes.byref(h)) != 0
    assert not h.value


@pytest.mark.parametrize("select", ["uniform", "learned"])
def test_guided_render_is_bit_reproducible(select, gpu_ctx):
    """Training records are compacted in slot order and every gradient sum is
    taken in a fixed order (no float atomics), and the learned lobe choice sums
    its statistics in fixed point, so a guided render -- the records, the Adam
    steps trained on them, the lobe choices, the guided samples and the film
    -- repeats bit for bit."""
    print("case guided reproducible", select, flush=True)
    spec = scenes.indirect_room_scene() if select == "uniform" else scenes.box_scene(seed=2, full=True)
    W, H = 48, 40
    stage = Stage(spec, ctx=gpu_ctx)
    out = []
    for _ in range(2):
        g = Guide(gpu_ctx, seed=5, batch=4096, steps_per_pass=4, train_paths=4096, warmup_steps=4, alpha=0.5,
                  guided_depth=2, lobe_select=select)
        f = Film(gpu_ctx, W, H)
        trace(stage, f, "guided", spp=24, max_depth=6, seed=61, batch=4, guide=g)
        out.append((f.readback(), g.get_params(), g.last_records().cpu().numpy(), g.steps()))
    (fa, pa, ra, sa), (fb, pb, rb, sb) = out
    assert sa == sb == 4 * 24 // 4
    assert ra.shape == rb.shape and ra.shape[0] > 100
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    assert np.array_equal(np.asarray(pa, np.float32).view(np.uint32), np.asarray(pb, np.float32).view(np.uint32))
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))


def test_guide_reduces_variance_under_indirect_light(gpu_ctx):
    print("case guided variance", flush=True)
    spec = scenes.indirect_room_scene()
    W, H = 32, 32
    stage = Stage(spec, ctx=gpu_ctx)
    g = Guide(gpu_ctx, batch=8192, steps_per_pass=8, train_paths=8192, warmup_steps=16, alpha=0.5)
    warm = Film(gpu_ctx, W, H)
    trace(stage, warm, "guided", spp=256, max_depth=6, seed=41, batch=8, guide=g)
    f1, f2 = Film(gpu_ctx, W, H), Film(gpu_ctx, W, H)
    a = _sample_radiance(stage, f1, "guided", 128, max_depth=6, seed=42, batch=8, guide=g).sum(2)
    b = _sample_radiance(stage, f2, "pt_nee", 128, max_depth=6, seed=43).sum(2)
    va, vb = a.var(0, ddof=1), b.var(0, ddof=1)
    lit = vb > 0
    ratio = va[lit].sum() / vb[lit].sum()
    print("variance ratio guided/plain %.3f" % ratio, flush=True)
    assert ratio < 0.7, ratio
    tot = abs(a.mean() - b.mean()) / np.sqrt(a.var() / a.size + b.var() / b.size)
    assert tot < 4.0, tot


def _guide_loss(g, recs):
    """The training loss on records: -(v / mean v) log q(direction), averaged."""
    import torch
    from gammaray_amd.api import mixture_pdf
    raw = recs[:, :8].contiguous()
    dirs = recs[:, 8:11].contiguous()
    v = recs[:, 11].clamp(min=0)
    mix = g.infer(raw)
    q = mixture_pdf(g.ctx, mix, dirs)
    w = (v / v.mean()).clamp(max=20.0)
    return float(-(w * torch.log(q.clamp(min=1e-30))).mean().item()), mix


@pytest.mark.parametrize("setting", ["consistent lobes", "all lobes", "bench"])
def test_guided_full_1080p_scene(setting, gpu_ctx):
    """Neural guiding on the full 1080p scene (100k-triangle mesh, 200 spheres,
    PT+NEE depth 8): alpha 0.3 at the first vertex, 8-spp batches, with either
    lobe set and the reference's lobe choice, and with the bench's setting
    (learned lobe choice, roulette 0.1, 4 Adam steps per batch).  Checks:
    * mixtures are finite and normalised;
    * training lowers the loss on the first batch's records, against an
      untrained network;
    * 16 guided spp agree with 16 plain spp (independent seeds) per 16 x 16
      tile within 3 sigma (binomial bound) and in total."""
    import torch
    print("case guided 1080p", setting, flush=True)
    spec = scenes.showcase_scene()
    W, H = 1920, 1080
    stage = Stage(spec, ctx=gpu_ctx)
    kw = dict(seed=7, learning_rate=5e-3, alpha=0.3, batch=16384, steps_per_pass=8, train_paths=65536,
              guided_depth=1, warmup_steps=8, lobes="all" if setting == "all lobes" else "consistent")
    if setting == "bench":
        kw.update(steps_per_pass=4, train_paths=1 << 19, lobe_select="learned", select_mix=0.4, rr_threshold=0.1)
    g = Guide(gpu_ctx, **kw)
    fresh = Guide(gpu_ctx, **kw)  # same initial weights, never trained
    warm = Film(gpu_ctx, W, H)
    trace(stage, warm, "guided", spp=8, max_depth=8, seed=51, batch=8, guide=g)
    r0 = g.last_records().clone()
    assert r0.shape[0] > 10000
    l_fresh, mix = _guide_loss(fresh, r0)
    l_one, _ = _guide_loss(g, r0)
    trace(stage, warm, "guided", spp=64, max_depth=8, seed=52, sample_offset=8, batch=8, guide=g)
    l_late, mix = _guide_loss(g, r0)
    print("loss on the first batch's records: untrained %.4f, 8 steps %.4f, %d steps %.4f"
          % (l_fresh, l_one, g.steps(), l_late), flush=True)
    assert l_one < l_fresh and l_late < l_one - 0.05, (l_fresh, l_one, l_late)
    m = mix.cpu().numpy().reshape(-1, 8, 5)
    assert np.isfinite(m).all()
    np.testing.assert_allclose(m[..., 0].sum(1), 1.0, atol=1e-3)
    assert (m[..., 1] > 0).all()
    a = _sample_radiance(stage, Film(gpu_ctx, W, H), "guided", 16, max_depth=8, seed=53, sample_offset=72,
                         batch=8, guide=g)
    b = _sample_radiance(stage, Film(gpu_ctx, W, H), "pt_nee", 16, max_depth=8, seed=54)
    assert np.isfinite(a).all()
    assert_three_sigma_blocks(a.sum(2), b.sum(2), W, H, 16, "guided 1080p tiles " + setting)


def _rel_mse_trimmed(img, ref, trim=1e-3):
    d = ((img - ref) ** 2 / (ref ** 2 + 1e-2)).ravel()
    d = d[np.isfinite(d)]
    return float(np.sort(d)[: int(len(d) * (1 - trim))].mean())


def test_guided_gains_at_equal_time(gpu_ctx):
    """The equal-time gain of guided rendering on the 1080p scene's geometry
    and camera at 960 x 540: 64 plain spp and 64 guided spp (the bench's guide
    setting, trained while rendering; both in 8-spp batches) against a 1024-spp
    plain render, error = relative MSE with the 0.1 % largest pixel errors
    dropped (the bench's measure).  Guided error times guided time must be
    below 0.8 of plain's (the bench measures about 0.5 at 1080p)."""
    import time
    print("case guided equal time", flush=True)
    spec = scenes.showcase_scene()
    W, H = 960, 540
    stage = Stage(spec, ctx=gpu_ctx)

    def image(f):
        px = f.readback()
        return px[..., :3] / px[..., 3:4]

    def timed(fn):
        gpu_ctx.synchronize()
        t = time.perf_counter()
        fn()
        gpu_ctx.synchronize()
        return time.perf_counter() - t

    ref = Film(gpu_ctx, W, H)
    trace(stage, ref, "pt_nee", spp=1024, max_depth=8, seed=1000)
    R = image(ref)
    warm = Film(gpu_ctx, W, H)
    trace(stage, warm, "pt_nee", spp=8, max_depth=8, seed=5)
    plain = Film(gpu_ctx, W, H)
    # both in 8-spp batches (the guide trains between batches; at 1080p plain's
    # default batch is 8 spp too)
    t_plain = timed(lambda: trace(stage, plain, "pt_nee", spp=64, max_depth=8, seed=7, batch=8))
    e_plain = _rel_mse_trimmed(image(plain), R)
    g = Guide(gpu_ctx, alpha=0.3, guided_depth=1, steps_per_pass=4, batch=16384, train_paths=1 << 19,
              warmup_steps=8, learning_rate=5e-3, lobe_select="learned", select_mix=0.4, rr_threshold=0.1)
    guided = Film(gpu_ctx, W, H)
    t_guided = 0.0
    for k in range(0, 64, 8):
        t_guided += timed(lambda: trace(stage, guided, "guided", spp=8, max_depth=8, seed=9, sample_offset=k, batch=8,
                                        guide=g))
    e_guided = _rel_mse_trimmed(image(guided), R)
    ratio = (e_guided * t_guided) / (e_plain * t_plain)
    print("plain %.5f in %.4f s, guided %.5f in %.4f s: error x time ratio %.3f"
          % (e_plain, t_plain, e_guided, t_guided, ratio), flush=True)
    assert e_guided < 0.85 * e_plain, (e_guided, e_plain)
    assert ratio < 0.8, ratio
