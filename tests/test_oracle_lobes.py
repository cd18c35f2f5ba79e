"""CPU tests pinning the oracle's clearcoat, dielectric Fresnel and specular
transmission against published formulas and derived distributions.

The expected values come from the literature, not from the reference:
  * GTR1 (Burley 2012, "Physically-Based Shading at Disney", appendix B):
    D(theta_h) = (a^2 - 1) / (pi ln(a^2) (1 + (a^2 - 1) cos^2 theta_h)),
    a = lerp(0.1, 0.001, gloss); its normalisation
    int D cos theta_h dw_h = 1, and its partial integral over
    cos^2 theta_h >= x0, 1 - ln(1 + (a^2 - 1) x0) / ln(a^2).
  * The reference's GTR1 sampler (src/material/util.rs:291-303) sets
    cos^2 theta_h = clamp(1 - (a^2)^(1 - u) / (1 - a^2)) -- a different
    precedence from Burley's (1 - (a^2)^(1 - u)) / (1 - a^2) (Q21).  With
    v = 1 - u uniform, P(cos^2 theta_h <= x) = ln((1 - x)(1 - a^2)) / ln(a^2),
    derived below; a chi-square test pins the sampler to it.
  * Dielectric Fresnel: the unpolarised Fresnel equations (Hecht, Optics,
    ch. 4): F = (r_s^2 + r_p^2) / 2, F(normal) = ((eta - 1) / (eta + 1))^2.
  * GGX: D = a^2 / (pi (cos^2 (a^2 - 1) + 1)^2) (Walter et al. 2007, eq. 33)
    and Smith's Lambda = (sqrt(1 + a^2 tan^2) - 1) / 2 (Heitz 2014).
  * Microfacet transmission (Walter et al. 2007, eqs. 16-21; PBRT 3e
    section 8.4.4): half vector h = normalize(i + eta o), Snell's law
    for refraction through h, pdf = D G1 |i.h| / |cos i| *
    eta^2 |o.h| / (i.h + eta o.h)^2, and the radiance/importance factor
    eta^2.  The reference's pdf does not normalise h when its z is positive
    (src/material/lobes.rs:337-339, Q22).  In the x-z plane, with
    |i + eta o| >= 1 and its z <= 1, the derivation in that test shows the
    reported pdf is the Walter/PBRT pdf with D evaluated at cos theta_h =
    z of the unnormalised vector.
"""
import math

import numpy as np
import pytest
from scipy import stats

from gammaray_amd import scenes
from oracle import oracle as O


def _scene(rec):
    s = scenes.SceneSpec()
    s.add_sphere((0, 0, 0), 1.0, rec)
    return O.OracleScene(s)


def _items(wi, wo, u=None, c2l=1.0):
    wi = np.atleast_2d(np.asarray(wi, np.float32))
    wo = np.atleast_2d(np.asarray(wo, np.float32))
    n = max(len(wi), len(wo))
    it = np.zeros((n, 24), np.float32)
    it[:, 0:3] = wi
    it[:, 3:6] = wo
    it[:, 6:9] = [0, 0, 1]     # shading normal
    it[:, 9:12] = [1, 0, 0]    # tangent
    it[:, 12:15] = [0, 1, 0]   # binormal
    it[:, 15:18] = [0, 0, 1]   # geometric normal
    if u is not None:
        it[:, 18] = u[:, 0].view(np.float32)
        it[:, 19:21] = u[:, 1:3].view(np.float32)
    it[:, 21] = c2l
    return it


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _query(o, wi, wo, u=None, c2l=1.0):
    return O.material_query(o, 0, _items(wi, wo, u, c2l))


# ---- GTR1 clearcoat --------------------------------------------------------
def _clearcoat_only(cc=1.0, gloss=0.5):
    # metallic 1: no diffuse, no transmission; ior 1: no specular lobe
    return _scene(scenes.disney(metallic=1.0, ior=1.0, clearcoat=cc, clearcoat_gloss=gloss))


def _gtr1_alpha(gloss):
    # the material stores a as an f32 (D ~ 1 / a^2 amplifies its rounding)
    f = np.float32
    return float(f(0.1) + f(gloss) * (f(0.001) - f(0.1)))


def _gtr1_d(c, a):
    a2 = a * a
    return (a2 - 1.0) / (math.pi * math.log(a2) * (1.0 + (a2 - 1.0) * c * c))


@pytest.mark.parametrize("gloss", [0.0, 0.5, 0.9])
def test_gtr1_peak_and_pdf_at_normal_incidence(gloss):
    o = _clearcoat_only(gloss=gloss)
    z = [0, 0, 1]
    q = _query(o, z, z)[0]
    assert q[18] == 1  # one lobe
    a = _gtr1_alpha(gloss)
    d1 = _gtr1_d(1.0, a)
    # pdf = D(h) / (4 i.h) with h = i = n: at normal incidence the missing
    # cos theta_h (Q21) is 1, so the value is the published D(0 degrees) / 4.
    # In f32, 1 + (a^2 - 1) cos^2 cancels to within 2^-24 of a^2: a relative
    # error up to 2^-24 / a^2.
    tol = max(2e-5, 4.0 * 2.0 ** -24 / (a * a))
    assert q[3] == pytest.approx(d1 / 4.0, rel=tol)


def _schlick(c, r0=0.04):
    return r0 + (1.0 - r0) * (1.0 - c) ** 5


@pytest.mark.parametrize("gloss", [0.0, 0.5, 0.9])
@pytest.mark.parametrize("theta_h", [30.0, 50.0, 60.0])
def test_gtr1_shape_off_normal(gloss, theta_h):
    """Two pairs with the same |cos i| = |cos o| = 0.45.  Pair A is the mirror
    pair about h = n.  Pair B is symmetric about h tilted by theta_h, with i and o
    leaving h out of the plane of tilt: i, o = cb h +- sqrt(1 - cb^2) y with
    cb = 0.45 / cos(theta_h), so i.z = o.z = 0.45.  The masking term depends only on
    |cos i| and |cos o|.  So f_B / f_A = F(o.h_B) D(theta_h) / (F(o.h_A) D(0))
    with the Schlick F (r0 = 0.04).  Burley's D gives
    D(theta_h) / D(0) = a^2 / (1 + (a^2 - 1) cos^2 theta_h)."""
    o = _clearcoat_only(gloss=gloss)
    ct = math.cos(math.radians(theta_h))
    ci = 0.45
    if ci / ct > 1.0:
        pytest.skip("no pair with this |cos i|")
    cb = ci / ct                      # o.h of pair B
    sb = math.sqrt(1.0 - cb * cb)
    h = np.array([math.sqrt(1.0 - ct * ct), 0.0, ct])
    y = np.array([0.0, 1.0, 0.0])
    i_b, o_b = cb * h + sb * y, cb * h - sb * y
    sa = math.sqrt(1.0 - ci * ci)
    f_a = _query(o, [sa, 0, ci], [-sa, 0, ci])[0, 0]
    f_b = _query(o, i_b, o_b)[0, 0]
    a = _gtr1_alpha(gloss)
    want = a * a / (1.0 + (a * a - 1.0) * ct * ct) * _schlick(cb) / _schlick(ci)
    tol = max(2e-4, 8.0 * 2.0 ** -24 / (a * a))
    assert f_b / f_a == pytest.approx(want, rel=tol)


def _smith_g_ggx(c, alpha_g):
    """Burley's BRDF-explorer smithG_GGX(NdotV, alphaG)."""
    a, b = alpha_g * alpha_g, c * c
    return 1.0 / (c + math.sqrt(a + b - a * b))


@pytest.mark.parametrize("cc,gloss", [(1.0, 0.3), (0.4, 0.8)])
def test_gtr1_value_at_mirror_pairs(cc, gloss):
    """For mirror pairs (h = n, o.h = cos theta), f = (0.25 cc) F D(0) G / (4 cos^2 theta).
    F is Schlick with r0 = 0.04 (the Disney clearcoat's ior 1.5) and D is
    Burley's GTR1.  Per Q21, G combines Burley's smithG_GGX at alpha_g = 0.25
    the way Smith's Lambda combines: G = 1 / (1 + g(cos i) + g(cos o))."""
    o = _clearcoat_only(cc=cc, gloss=gloss)
    a = _gtr1_alpha(gloss)
    d0 = _gtr1_d(1.0, a)
    tol = max(3e-5, 4.0 * 2.0 ** -24 / (a * a))
    for th in (0.0, 25.0, 50.0, 70.0):
        c, s = math.cos(math.radians(th)), math.sin(math.radians(th))
        f = _query(o, [s, 0, c], [-s, 0, c])[0, 0]
        g = 1.0 / (1.0 + 2.0 * _smith_g_ggx(c, 0.25))
        want = 0.25 * cc * _schlick(c) * d0 * g / (4.0 * c * c)
        assert f == pytest.approx(want, rel=tol), th


@pytest.mark.parametrize("gloss", [0.1, 0.5])
def test_gtr1_partial_normalisation(gloss):
    """For i = n the reflected o has cos theta_h = i.h and dw_o = 4 (o.h) dw_h.
    So the integral over o of pdf(o) * cos theta_h equals the integral of
    D(h) cos theta_h over the half vectors with cos^2 theta_h >= 1/2.  In
    closed form that is 1 - ln((1 + a^2) / 2) / ln(a^2)."""
    o = _clearcoat_only(gloss=gloss)
    a = _gtr1_alpha(gloss)
    # quadrature over x = cos^2 theta_h in (1/2, 1), densely near 1 where D peaks
    t = np.linspace(0.0, 1.0, 4001)
    x = 1.0 - 0.5 * (a * a * 1e-3 / 0.5) ** t  # 1 - x log-spaced from 1/2 to a^2 / 2000
    x = np.concatenate([[0.5], x[1:], [1.0]])
    c = np.sqrt(x)
    h = np.stack([np.sqrt(1 - x), np.zeros_like(x), c], 1)
    wo = 2 * c[:, None] * h - np.array([0, 0, 1.0])
    pdf = _query(o, np.tile([0, 0, 1.0], (len(x), 1)), wo)[:, 3].astype(np.float64)
    # dw_o = 4 (o.h) dw_h and dw_h = dphi dx / (2 cos theta_h): the o-integral of
    # pdf * cos theta_h is the x-integral of pdf * 4 c * pi
    integrand = pdf * 4.0 * c * math.pi
    got = np.trapezoid(integrand, x)
    want = 1.0 - math.log((1.0 + a * a) / 2.0) / math.log(a * a)
    assert got == pytest.approx(want, rel=2e-3)


def _gtr1_sampler_cdf(x, a):
    # P(cos^2 theta_h <= x) = ln((1 - x)(1 - a^2)) / ln(a^2), derived from
    # x = 1 - A / (1 - a^2), A = (a^2)^v, v uniform on (0, 1].
    a2 = a * a
    return np.log((1.0 - x) * (1.0 - a2)) / math.log(a2)


@pytest.mark.parametrize("gloss", [0.2, 0.7])
def test_gtr1_sampler_distribution(gloss):
    o = _clearcoat_only(gloss=gloss)
    a = _gtr1_alpha(gloss)
    n = 200000
    rng = np.random.default_rng(5)
    u = np.zeros((n, 3), np.uint32)
    u[:, 1:3] = rng.random((n, 2)).astype(np.float32).view(np.uint32)
    q = _query(o, np.tile([0, 0, 1.0], (n, 1)), np.tile([0, 0, 1.0], (n, 1)), u)
    ok = q[:, 17] != 0
    # half vectors with cos^2 theta_h <= 1/2 reflect below the surface and are
    # rejected: the acceptance rate is 1 - F(1/2)
    acc = ok.mean()
    p_acc = 1.0 - _gtr1_sampler_cdf(0.5, a)
    assert abs(acc - p_acc) < 5 * math.sqrt(p_acc * (1 - p_acc) / n)
    wo = q[ok, 13:16].astype(np.float64)
    x = (wo[:, 2] + 1.0) / 2.0  # o = 2 (i.h) h - i at i = n: o.z = 2 cos^2 theta_h - 1
    # equal-probability bins of the derived distribution, conditioned on x > 1/2
    k = 40
    f_half = _gtr1_sampler_cdf(0.5, a)
    levels = f_half + (1.0 - f_half) * np.arange(1, k) / k
    edges = 1.0 - np.exp(levels * math.log(a * a)) / (1.0 - a * a)
    counts = np.bincount(np.searchsorted(edges, x), minlength=k)
    chi2 = stats.chisquare(counts).pvalue
    assert chi2 > 1e-4, chi2
    # Burley's precedence, cos^2 = (1 - (a^2)^v) / (1 - a^2), has the CDF
    # ln(1 - x (1 - a^2)) / ln(a^2); the same samples reject it on these bins.
    a2 = a * a
    fb = np.log(1.0 - np.concatenate([[0.5], edges, [1.0]]) * (1.0 - a2)) / math.log(a2)
    pb = np.diff(fb) / (1.0 - fb[0])
    assert stats.chisquare(counts, pb * counts.sum()).pvalue < 1e-6
    # azimuth uniform
    phi = np.arctan2(wo[:, 1], wo[:, 0])
    cnt_phi = np.histogram(phi, bins=24, range=(-math.pi, math.pi))[0]
    assert stats.chisquare(cnt_phi).pvalue > 1e-4
    # Seen from below (i = -n), the half vector is flipped into i's hemisphere,
    # h -> -h.  The same draws then give the point-reflected direction -o with
    # the same positive pdf, since D(h) / (4 i.h) needs i.h > 0.
    m = 5000
    below = _query(o, np.tile([0, 0, -1.0], (m, 1)), np.tile([0, 0, -1.0], (m, 1)), u[:m])
    above = q[:m]
    assert np.array_equal(below[:, 17] != 0, above[:, 17] != 0)
    live = above[:, 17] != 0
    np.testing.assert_allclose(below[live, 13:16], -above[live, 13:16], atol=1e-6)
    assert np.all(below[live, 16] > 0)
    np.testing.assert_allclose(below[live, 16], above[live, 16], rtol=1e-6)


# ---- dielectric Fresnel ----------------------------------------------------
def _fresnel(cos_i, eta_i, eta_t):
    """Unpolarised Fresnel reflectance (Hecht): 1 under total internal reflection."""
    sin_t = eta_i / eta_t * math.sqrt(max(0.0, 1.0 - cos_i * cos_i))
    if sin_t >= 1.0:
        return 1.0
    cos_t = math.sqrt(1.0 - sin_t * sin_t)
    rs = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t)
    rp = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t)
    return 0.5 * (rs * rs + rp * rp)


def _ggx_d(c, a):
    c2 = c * c
    return a * a / (math.pi * (c2 * (a * a - 1.0) + 1.0) ** 2)


def _ggx_lambda(c, a):
    t2 = (1.0 - c * c) / (c * c)
    return 0.5 * (math.sqrt(1.0 + a * a * t2) - 1.0)


def _specular_dielectric(ior, roughness):
    # metallic 0, specular_tint 0: DisneyFresnel = F_dielectric; specular_trans 1
    # removes the diffuse lobe (its transmission lobe has no reflection f)
    return _scene(scenes.disney(roughness=roughness, ior=ior, specular_trans=1.0))


@pytest.mark.parametrize("ior", [1.005, 1.2, 1.5, 2.4])
def test_dielectric_fresnel_at_normal_incidence(ior):
    r = 0.6
    o = _specular_dielectric(ior, r)
    z = [0, 0, 1]
    q = _query(o, z, z)[0]
    a = r * r
    # specular f at i = o = n: F(1) D(n) G / 4 with D(n) = 1 / (pi a^2) and G = 1
    f0 = q[0] * 4.0 * math.pi * a * a
    eta = max(ior, 1.01)  # DisneySpecularRefl clamps ior to >= 1.01
    assert f0 == pytest.approx(((eta - 1) / (eta + 1)) ** 2, rel=1e-4)


@pytest.mark.parametrize("ior", [1.3, 1.5, 2.0])
def test_dielectric_fresnel_oblique(ior):
    """Mirror pairs (h = n, o.h = cos theta): f = F(cos) D(n) G(i, o) / (4 cos^2)."""
    r = 0.5
    a = r * r
    o = _specular_dielectric(ior, r)
    for th in (10.0, 40.0, 65.0, 80.0):
        c, s = math.cos(math.radians(th)), math.sin(math.radians(th))
        f = _query(o, [s, 0, c], [-s, 0, c])[0, 0]
        g = 1.0 / (1.0 + 2.0 * _ggx_lambda(c, a))
        fr = f * 4.0 * c * c / (_ggx_d(1.0, a) * g)
        assert fr == pytest.approx(_fresnel(c, 1.0, ior), rel=2e-4), th


# ---- specular transmission -------------------------------------------------
def _transmissive(ior=1.5, roughness=0.3, base=(0.64, 0.36, 0.81)):
    # lobes: GGX specular (ior > 1) then transmission (colour sqrt(base))
    return _scene(scenes.disney(base_color=base, roughness=roughness, ior=ior, specular_trans=1.0))


def test_transmission_snell_limit():
    """At the minimum roughness (a = 0.001) sampled half vectors lie within a
    few 1e-3 of the normal, so the sampled direction is Snell's refraction;
    inside rays beyond the critical angle refract to nothing (TIR)."""
    ior = 1.5
    o = _transmissive(ior=ior, roughness=0.0)
    rng = np.random.default_rng(8)
    n = 400
    u = np.zeros((n, 3), np.uint32)
    u[:, 0] = 0xF0000000  # the second of two lobes: transmission
    u[:, 1:3] = rng.uniform(0.05, 0.95, (n, 2)).astype(np.float32).view(np.uint32)
    th = np.radians(rng.uniform(1.0, 85.0, n))
    phi = rng.uniform(0, 2 * math.pi, n)
    wi = np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], 1)
    # entering from outside
    q = _query(o, wi, wi, u)
    assert np.all(q[:, 17] == 4 + 16)  # GLOSSY | TRANSMISSION
    wo = q[:, 13:16]
    sin_o = np.sin(th) / ior
    want = np.stack([-np.sin(th) * np.cos(phi) / ior, -np.sin(th) * np.sin(phi) / ior,
                     -np.sqrt(1 - sin_o ** 2)], 1)
    np.testing.assert_allclose(wo, want, atol=0.02)
    # exiting from inside: sin(out) = ior sin(in); beyond asin(1 / ior) nothing
    wi_in = wi * np.array([1, 1, -1.0])
    q = _query(o, wi_in, wi_in, u)
    crit = np.sin(th) * ior >= 1.0
    margin = np.abs(np.sin(th) * ior - 1.0) > 0.02
    assert np.all(q[crit & margin, 17] == 0)
    ok = ~crit & margin
    assert np.all(q[ok, 17] == 4 + 16)
    s_out = np.sin(th[ok]) * ior
    want = np.stack([-s_out * np.cos(phi[ok]), -s_out * np.sin(phi[ok]), np.sqrt(1 - s_out ** 2)], 1)
    np.testing.assert_allclose(q[ok, 13:16], want, atol=0.03)


def _true_trans_pdf(i, o, eta, a):
    hu = i + eta * o
    h = _unit(hu) * (1.0 if hu[2] > 0 else -1.0)
    d = _ggx_d(h[2], a)
    g1 = 1.0 / (1.0 + _ggx_lambda(abs(i[2]), a))
    sd = i @ h + eta * (o @ h)
    return d * g1 * abs(i @ h) / abs(i[2]) * eta * eta * abs(o @ h) / (sd * sd), hu


def _trans_pairs(in_plane=False):
    out = []
    rng = np.random.default_rng(12)
    for _ in range(400):
        i = _unit(rng.normal(size=3))
        o = _unit(rng.normal(size=3))
        if in_plane:
            i[1] = o[1] = 0.0
            i, o = _unit(i), _unit(o)
        if abs(i[2]) < 0.05 or abs(o[2]) < 0.05:
            continue
        o[2] = -math.copysign(abs(o[2]), i[2])  # opposite hemispheres
        out.append((i, o))
    return out


def test_transmission_pdf_and_its_unnormalised_half_vector():
    """Pairs whose half vector i + eta o points below the surface use the
    normalised h: the Walter/PBRT pdf exactly.  Pairs whose i + eta o points
    up keep it unnormalised (Q22).  D then works from spherical coordinates
    of a non-unit vector: sin^2 theta = 1 - z^2 and cos phi = clamp(x / sin theta).
    For pairs in the x-z plane with |hu| >= 1 and z <= 1, cos phi still clamps
    to +-1, as the unit vector's is.  With tan^2 = (1 - z^2) / z^2, the GGX
    D = 1 / (pi a^2 cos^4 (1 + tan^2 / a^2)^2) equals Walter's
    a^2 / (pi (z^2 (a^2 - 1) + 1)^2) at cos = z.  The factor |hu| in |i.h|
    cancels the 1 / |hu| of the Jacobian.  So the reported pdf is the
    Walter/PBRT pdf with D evaluated at cos theta_h = hu.z."""
    ior, r = 1.5, 0.4
    a = r * r
    o = _transmissive(ior=ior, roughness=r)
    n_pos = n_neg = 0
    for in_plane in (False, True):
        pairs = _trans_pairs(in_plane)
        wi = np.array([p[0] for p in pairs], np.float32)
        wo = np.array([p[1] for p in pairs], np.float32)
        q = _query(o, wi, wo)
        for k in range(len(pairs)):
            i = wi[k].astype(np.float64)
            w = wo[k].astype(np.float64)
            eta = ior if i[2] > 0 else 1.0 / ior
            p, hu = _true_trans_pdf(i, w, eta, a)
            if hu[2] > 0:
                if not in_plane or np.linalg.norm(hu) < 1.0 or hu[2] > 1.0:
                    continue
                p *= _ggx_d(hu[2], a) / _ggx_d(_unit(hu)[2], a)   # Q22
                n_pos += 1
            else:
                n_neg += 1
            # pdf_world averages the two lobes; the specular pdf is 0 across the
            # surface.  f32 dot products near zero lose relative precision.
            hn = _unit(hu)
            tol = 5e-4 + 8 * 2.0 ** -24 * (1 / abs(w @ hn) + 1 / abs(i @ hn))
            assert 2.0 * q[k, 3] == pytest.approx(p, rel=tol, abs=1e-9), (k, hu)
    assert n_pos > 10 and n_neg > 50, (n_pos, n_neg)


def test_transmission_zero_in_the_same_hemisphere():
    # ior 1: no specular lobe, so the material is the transmission lobe alone
    # (its own ior clamped to 1.01)
    o = _scene(scenes.disney(roughness=0.5, ior=1.0, specular_trans=1.0))
    pairs = _trans_pairs()
    wi = np.array([p[0] for p in pairs], np.float32)
    wo = np.array([p[1] for p in pairs], np.float32)
    across = _query(o, wi, wo)
    assert across[0, 18] == 1
    assert np.all(across[:, 3] > 0) and (across[:, 0] > 0).mean() > 0.5
    same = _query(o, wi, wo * np.array([1, 1, -1], np.float32))
    assert np.all(same[:, 0:4] == 0)


def test_transmission_importance_factor():
    """Radiance (camera -> light) carries factor 1, importance eta^2, with
    eta = ior entering and 1 / ior leaving (src/material/lobes.rs:310)."""
    ior = 1.5
    o = _transmissive(ior=ior, roughness=0.5)
    pairs = _trans_pairs()
    wi = np.array([p[0] for p in pairs], np.float32)
    wo = np.array([p[1] for p in pairs], np.float32)
    f_rad = _query(o, wi, wo, c2l=1.0)[:, 0:3]
    f_imp = _query(o, wi, wo, c2l=0.0)[:, 0:3]
    eta = np.where(wi[:, 2] > 0, ior, 1.0 / ior)
    live = f_rad[:, 0] > 0
    assert live.sum() > 50
    np.testing.assert_allclose(f_imp[live] / f_rad[live], np.repeat((eta * eta)[live, None], 3, 1), rtol=1e-5)


def test_transmission_f_matches_walter():
    """f = color (1 - F(o.h)) D G |o.h| |i.h| / (|cos o| |cos i| (i.h + eta o.h)^2)
    for radiance (camera -> light, factor 1), colour = specular_trans * sqrt(base)."""
    ior, r = 1.5, 0.4
    a = r * r
    base = np.array([0.64, 0.36, 0.81])
    o = _transmissive(ior=ior, roughness=r, base=tuple(base))
    pairs = _trans_pairs()
    wi = np.array([p[0] for p in pairs], np.float32)
    wo = np.array([p[1] for p in pairs], np.float32)
    q = _query(o, wi, wo, c2l=1.0)
    checked = 0
    for k in range(len(pairs)):
        i = wi[k].astype(np.float64)
        w = wo[k].astype(np.float64)
        eta = ior if i[2] > 0 else 1.0 / ior
        hu = i + eta * w
        h = _unit(hu) * (1.0 if hu[2] > 0 else -1.0)
        oh, ih = w @ h, i @ h
        # Fresnel of the refracted side: o.h > 0 means o leaves through the
        # outside (eta 1 -> ior), o.h < 0 through the inside
        fr = _fresnel(oh, 1.0, ior) if oh > 0 else _fresnel(-oh, ior, 1.0)
        g = 1.0 / (1.0 + _ggx_lambda(abs(i[2]), a) + _ggx_lambda(abs(w[2]), a))
        sd = ih + eta * oh
        want = np.sqrt(base) * (1.0 - fr) * abs(_ggx_d(h[2], a) * g * abs(oh) * abs(ih) /
                                                  (w[2] * i[2] * sd * sd))
        tol = 5e-4 + 8 * 2.0 ** -24 * (1 / abs(oh) + 1 / abs(ih) + (abs(ih) + eta * abs(oh)) / abs(sd))
        np.testing.assert_allclose(q[k, 0:3], want, rtol=tol, atol=1e-7)
        checked += fr < 1.0
    assert checked > 50


# ---- the density of the draws ------------------------------------------------
# The guided estimator weighs a direction by the density the lobes' samplers
# realise there (material_query column 19).  Pinned against the oracle's own
# sample_world: the draws of one incoming direction, binned on the sphere
# (equal-area cells in z and azimuth), are compared by chi-square with the
# density's integral over each cell (midpoint rule on a finer grid), and the
# draws that return no direction with one minus the density's total.  A
# density that used the reported pdfs (GTR1 without cos theta_h, the
# transmission pdf's unnormalised half vector), Burley's GTR1 precedence,
# dropped the reflection Jacobian, or left out the transmission draws that
# land in the incoming hemisphere fails these bins.
def _density_chi2(o, wi, n_draws=300000, kz=24, kphi=24, sub=24, seed=11):
    rng = np.random.default_rng(seed)
    u = np.zeros((n_draws, 3), np.uint32)
    u[:, 0] = rng.integers(0, 1 << 32, n_draws, dtype=np.uint64).astype(np.uint32)
    u[:, 1:3] = rng.random((n_draws, 2)).astype(np.float32).view(np.uint32)
    wi = _unit(wi).astype(np.float32)
    q = _query(o, np.tile(wi, (n_draws, 1)), np.tile([0, 0, 1.0], (n_draws, 1)), u)
    d = q[:, 13:16].astype(np.float64)
    drawn = np.any(d != 0.0, axis=1)
    z = np.clip(d[drawn, 2], -1.0, 1.0)
    phi = np.arctan2(d[drawn, 1], d[drawn, 0])
    iz = np.minimum(((z + 1.0) * 0.5 * kz).astype(int), kz - 1)
    ip = np.minimum(((phi + math.pi) / (2 * math.pi) * kphi).astype(int), kphi - 1)
    counts = np.bincount(iz * kphi + ip, minlength=kz * kphi)
    nz, nphi = kz * sub, kphi * sub
    zc = -1.0 + (np.arange(nz) + 0.5) * (2.0 / nz)
    pc = -math.pi + (np.arange(nphi) + 0.5) * (2.0 * math.pi / nphi)
    zz, pp = np.meshgrid(zc, pc, indexing="ij")
    ss = np.sqrt(1.0 - zz * zz)
    dirs = np.stack([ss * np.cos(pp), ss * np.sin(pp), zz], -1).reshape(-1, 3)
    dens = _query(o, np.tile(wi, (len(dirs), 1)), dirs)[:, 19].astype(np.float64)
    assert np.all(dens >= 0.0)
    cell = (2.0 / nz) * (2.0 * math.pi / nphi)
    mass = dens.reshape(kz, sub, kphi, sub).sum(axis=(1, 3)).ravel() * cell
    total = mass.sum()
    assert total <= 1.0 + 2e-3, total
    exp = np.append(mass, max(0.0, 1.0 - total)) * n_draws
    obs = np.append(counts, n_draws - drawn.sum()).astype(np.float64)
    small = exp < 5.0
    exp = np.append(exp[~small], exp[small].sum())
    obs = np.append(obs[~small], obs[small].sum())
    chi = ((obs - exp) ** 2 / np.maximum(exp, 1e-12)).sum()
    p = stats.chi2.sf(chi, len(exp) - 1)
    return p, total, drawn.mean(), d[drawn], wi


@pytest.mark.parametrize("gloss", [0.0, 0.3])
@pytest.mark.parametrize("wz", [0.95, 0.3, -0.6])
def test_clearcoat_draw_density(gloss, wz):
    o = _clearcoat_only(gloss=gloss)
    p, total, acc, _, _ = _density_chi2(o, [math.sqrt(1 - wz * wz), 0.0, wz])
    assert p > 1e-4, (p, total, acc)
    assert abs(total - acc) < 5e-3, (total, acc)


@pytest.mark.parametrize("wz", [0.9, 0.25, -0.45, -0.9])
@pytest.mark.parametrize("roughness", [0.35, 0.5])
def test_transmission_draw_density(wz, roughness):
    """Two lobes, GGX reflection and transmission: the mean of their densities.
    The GGX sampler inverts the slope distribution with PBRT's rational fit
    (src/material/util.rs:194-223), so its draws follow the visible-normal
    density only as closely as that fit; at roughness 0.7 the steepest slopes,
    which refract to the edge of the transmission cone, already differ by a
    few hundred in a million draws there, and these bins stay below that."""
    o = _transmissive(ior=1.5, roughness=roughness)
    p, total, acc, d, wi = _density_chi2(o, [math.sqrt(1 - wz * wz), 0.0, wz])
    assert p > 1e-4, (p, total, acc)
    assert abs(total - acc) < 5e-3, (total, acc)


def test_transmission_draws_reach_the_incoming_hemisphere():
    """Leaving the medium at grazing incidence, o = -eta' i + (eta' i.h - cos t) h
    (eta' = ior) has o.z = eta' |i.z| - (eta' i.h - cos t) |h.z|, which a tilted
    h makes negative: some refracted draws land on the incoming side (entering,
    both terms keep o.z < 0).  The density covers them, although the reported
    transmission pdf is zero there: with the GGX lobe's reported pdf
    p_r = 2 pdf_world on that side, the density (p_r + s_t) / 2 exceeds
    pdf_world by s_t / 2 > 0."""
    o = _transmissive(ior=1.5, roughness=0.5)
    wz = -0.15
    p, total, acc, d, wi = _density_chi2(o, [math.sqrt(1 - wz * wz), 0.0, wz])
    assert p > 1e-4, (p, total, acc)
    n = 200000
    rng = np.random.default_rng(3)
    u = np.zeros((n, 3), np.uint32)
    u[:, 0] = (1 << 31) + rng.integers(0, 1 << 31, n, dtype=np.uint64).astype(np.uint32)  # lobe 1 of 2
    u[:, 1:3] = rng.random((n, 2)).astype(np.float32).view(np.uint32)
    q = _query(o, np.tile(wi, (n, 1)), np.tile([0, 0, 1.0], (n, 1)), u)
    dd = q[:, 13:16]
    side = dd[:, 2] * wi[2] > 0
    assert side.mean() > 0.01, side.mean()
    r = _query(o, np.tile(wi, (int(side.sum()), 1)), dd[side])
    assert np.all(r[:, 19] > r[:, 3]), np.min(r[:, 19] - r[:, 3])


def test_full_disney_draw_density():
    rng = np.random.default_rng(17)
    for k in range(3):
        rec = scenes.random_disney(rng, full=True)
        rec[12] = max(rec[12], 0.3)              # clearcoat on
        rec[5] = 0.0                             # isotropic
        rec[4] = 0.3 + 0.2 * rec[4]              # roughness in [0.3, 0.5]
        o = _scene(rec)
        for wz in (0.8, -0.4):
            p, total, acc, _, _ = _density_chi2(o, [math.sqrt(1 - wz * wz), 0.0, wz], n_draws=200000)
            assert p > 1e-4, (k, wz, p, total, acc)
