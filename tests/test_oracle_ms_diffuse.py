"""Pins the oracle's microsurface walks of the rough diffuse material and the
Beckmann multiple-scattering dielectric evaluation (CU/rough_diffuse_shader.cu:
126-139, 187-216, 252-268; ORF/MicrofacetGGX.h:342-409, 519-611;
ORF/MicrofacetBeckmann.h:375-442, 479-527, 613-708), CPU only.

- White furnace: with rho = 1 the GGX diffuse walk's weight is 1 and every walk
  leaves upwards; the evaluation truncated after k events integrates over the
  sphere to the fraction of sampled walks that leave within k events
  (k = 1, 3, 100), so the full evaluation integrates to 1.
- First order, in closed form for both distributions: the first event's height
  h1 on a ray entering from above along -w_i has C1(h1) = V^(1 / (1 + L_i)),
  V uniform, so the masking C1(h1)^L_o of the exit direction averages to
  (1 + L_i) / (1 + L_i + L_o) (the height-correlated Smith ratio).  The
  evaluations truncated after one event therefore integrate to
    diffuse:    E[(1 + L_i) / (1 + L_i + L_o)], m ~ D_wi, w_o ~ cos about m;
    conductor:  E[F(w_i, h) |w_i.h| (1 + L_i) / (1 + L_i + L_o)], h ~ D_wi,
                w_o the mirror of w_i about h (the reference's phase is
                D_wi(h) / 4 without the 1 / |w_i.h| of D_wi / (4 |w_i.h|));
    dielectric: E[F(w_i, h) (1 + L_i) / (1 + L_i + L_o)] over the reflected
                exits (Beckmann; the GGX one is pinned against its sampler).
  L is the Smith Lambda written here from its closed forms (GGX
  (-1 + sqrt(1 + a^2 tan^2)) / 2; Beckmann (erf(b) - 1) / 2 +
  exp(-b^2) / (2 b sqrt(pi)), b = 1 / (a tan)), m and h come from the visible-
  normal samplers pinned elsewhere.  A dropped masking factor, a wrong height
  sample or a swapped eta / 1 / eta moves these integrals by far more than
  their Monte Carlo error.
- The Beckmann diffuse sampler leaves after one event with the height-averaged
  probability 1 / (1 + L(w_1)) (beckmann_sample_height's escape), not the
  height-correlated one, so its first-order fraction is
  E[1 / (1 + L(w_1))]; its evaluation beyond order 1 does not integrate to its
  sampler (the reference mixes the two masking forms), which the GGX walks'
  furnace pins in their stead (the walk code is shared).
- Smooth limit: at roughness 0.02 the diffuse walk is one Lambertian event:
  the evaluation is rho cos / pi on average (mean ratio within 0.3%) and the
  sampled exits are cosine-distributed with weight rho.
- Structure: the sampler's weight is rho^k after k events, and its exit lies in
  the hemisphere of its last visible normal.
"""
import math

import numpy as np
import pytest
from scipy.special import erf

from tests.oracle_api import Oracle

MIRROR = ((1.0, 1e4),) * 3
IOR = 1.5


@pytest.fixture(scope="module")
def api():
    return Oracle()


def _wi(theta):
    t = math.radians(theta)
    return np.array([math.sin(t), 0.0, math.cos(t)], np.float32)


def _lambda(w, alpha, dist):
    """Smith Lambda of unit directions w (rows) above the mean plane."""
    z = np.clip(w[:, 2].astype(np.float64), -1.0, 1.0)
    tan = np.sqrt(np.maximum(0.0, 1.0 - z * z)) / np.maximum(z, 1e-12)
    if dist == 1:
        return 0.5 * (-1.0 + np.sqrt(1.0 + (alpha * tan) ** 2))
    b = 1.0 / np.maximum(alpha * tan, 1e-12)
    return np.where(b > 30.0, 0.0, 0.5 * (erf(b) - 1.0) + np.exp(-b * b) / (2.0 * b * math.sqrt(math.pi)))


def _sphere(n, seed):
    v = np.random.default_rng(seed).normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _cos_about(m, rng):
    """Cosine-distributed unit directions about the unit rows of m."""
    n = len(m)
    u1, u2 = rng.random(n), rng.random(n)
    r, phi = np.sqrt(u1), 2.0 * math.pi * u2
    loc = np.stack([r * np.cos(phi), r * np.sin(phi), np.sqrt(1.0 - u1)], axis=1)
    a = np.where(np.abs(m[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(m, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(m, t)
    return loc[:, :1] * t + loc[:, 1:2] * b + loc[:, 2:] * m


def _vndf(api, wi, alpha, dist, n, seed):
    rng = np.random.default_rng(seed)
    z1, z2 = rng.random(n, dtype=np.float32), rng.random(n, dtype=np.float32)
    wis = np.repeat(wi[None, :], n, axis=0)
    f = api.ggx_sample_vndf if dist == 1 else api.beckmann_sample_vndf
    return f(wis, alpha, alpha, z1, z2).astype(np.float64)


def _integral(f, up, m):
    """Sphere integral (estimate, standard error) of f sampled at m uniform directions, upper half only."""
    g = np.where(up, f, 0.0) * 4.0 * math.pi
    return g.mean(), g.std() / math.sqrt(m)


def _fresnel(c, eta):
    st2 = eta * eta * (1.0 - c * c)
    ct = np.sqrt(np.maximum(0.0, 1.0 - st2))
    rs = (eta * c - ct) / (eta * c + ct)
    rp = (c - eta * ct) / (c + eta * ct)
    return np.where(st2 < 1.0, 0.5 * (rs * rs + rp * rp), 1.0)


@pytest.mark.parametrize("alpha,theta", [(0.5, 20.0), (0.9, 70.0)])
def test_white_ggx_diffuse_walk_furnace(api, alpha, theta):
    wi = _wi(theta)
    wo, wm, w, order = api.ms_diffuse_sample(wi, alpha, (1.0, 1.0, 1.0), 200000, dist=1, seed=5)
    assert np.all(wo[:, 2] > 0.0) and np.all(order >= 1) and np.all(order < 100)
    np.testing.assert_array_equal(w, 1.0)
    m = 400000
    v = _sphere(m, 3)
    up = v[:, 2] > 0
    for cap in (1, 3, 100):
        f = api.ms_diffuse_eval(wi, v, alpha, (1.0, 1.0, 1.0), dist=1, seed=9, cap=cap).astype(np.float64)
        assert np.all(f[~up] == 0.0)
        e, se = _integral(f[:, 0], up, m)
        s = (order <= cap).mean()
        assert abs(e - s) < max(4 * se, 1e-3), (cap, e, s, se)
        if cap == 100:
            assert abs(e - 1.0) < 4 * se + 1e-3, e


@pytest.mark.parametrize("dist", [1, 0])
@pytest.mark.parametrize("alpha,theta", [(0.5, 20.0), (0.9, 70.0)])
def test_first_order_evaluations_closed_form(api, dist, alpha, theta):
    wi = _wi(theta)
    li = _lambda(wi[None, :], alpha, dist)[0]
    n = 400000
    rng = np.random.default_rng(21)
    hm = _vndf(api, wi, alpha, dist, n, seed=17)
    m = 400000
    v = _sphere(m, 4)
    up = v[:, 2] > 0
    # diffuse
    wo = _cos_about(hm, rng)
    g = np.where(wo[:, 2] > 0, (1 + li) / (1 + li + _lambda(wo, alpha, dist)), 0.0)
    want, se_w = g.mean(), g.std() / math.sqrt(n)
    f = api.ms_diffuse_eval(wi, v, alpha, (1.0, 1.0, 1.0), dist=dist, seed=9, cap=1)[:, 0].astype(np.float64)
    e, se = _integral(f, up, m)
    assert abs(e - want) < 4 * math.hypot(se, se_w) + 1e-3, ("diffuse", e, want, se, se_w)
    # conductor: mirror, then the dielectric interface as a conductor of index 1 / IOR
    cih = hm @ wi.astype(np.float64)
    wr = 2.0 * cih[:, None] * hm - wi[None, :]
    gm = np.where(wr[:, 2] > 0, cih * (1 + li) / (1 + li + _lambda(wr, alpha, dist)), 0.0)
    for ior, F in ((MIRROR, 1.0), (((IOR, 0.0),) * 3, _fresnel(cih, 1.0 / IOR))):
        t = gm * F
        want, se_w = t.mean(), t.std() / math.sqrt(n)
        f = api.ms_conductor_eval(wi, v, ior, alpha, dist=dist, seed=9, cap=1)[:, 0].astype(np.float64)
        e, se = _integral(f, up, m)
        assert abs(e - want) < 4 * math.hypot(se, se_w) + 1e-3 * want, ("conductor", ior[0], e, want, se, se_w)
    if dist == 0:
        # Beckmann dielectric evaluation, reflected exits
        t = np.where(wr[:, 2] > 0, _fresnel(cih, 1.0 / IOR) * (1 + li) / (1 + li + _lambda(wr, alpha, dist)), 0.0)
        want, se_w = t.mean(), t.std() / math.sqrt(n)
        f = api.beckmann_ms_eval(wi, v, 1.0 / IOR, alpha, seed=9, cap=1).astype(np.float64)
        e, se = _integral(f, up, m)
        assert abs(e - want) < 4 * math.hypot(se, se_w) + 1e-3 * want, ("dielectric", e, want, se, se_w)


@pytest.mark.parametrize("dist", [1, 0])
@pytest.mark.parametrize("alpha,theta", [(0.5, 20.0), (0.9, 70.0)])
def test_diffuse_sampler_first_order_fraction(api, dist, alpha, theta):
    wi = _wi(theta)
    wo_s, _, _, order = api.ms_diffuse_sample(wi, alpha, (1.0, 1.0, 1.0), 300000, dist=dist, seed=5)
    n = 400000
    hm = _vndf(api, wi, alpha, dist, n, seed=17)
    w1 = _cos_about(hm, np.random.default_rng(22))
    if dist == 1:
        li = _lambda(wi[None, :], alpha, dist)[0]
        g = np.where(w1[:, 2] > 0, (1 + li) / (1 + li + _lambda(w1, alpha, dist)), 0.0)
    else:
        g = np.where(w1[:, 2] > 0, 1.0 / (1.0 + _lambda(w1, alpha, dist)), 0.0)
    want, se_w = g.mean(), g.std() / math.sqrt(n)
    got = (order == 1).mean()
    se = math.sqrt(got * (1 - got) / len(order))
    assert abs(got - want) < 4 * math.hypot(se, se_w) + 1e-3, (got, want)


@pytest.mark.parametrize("dist", [1, 0])
def test_smooth_limit_is_lambert(api, dist):
    rho = np.array([0.7, 0.5, 0.3])
    for theta in (10.0, 50.0):
        wi = _wi(theta)
        wo, wm, w, order = api.ms_diffuse_sample(wi, 0.02, rho, 100000, dist=dist, seed=7)
        one = order == 1
        assert one.mean() > 0.99
        np.testing.assert_allclose(w[one], np.broadcast_to(rho, w[one].shape), rtol=1e-6)
        c2 = np.sort(wo[one, 2].astype(np.float64) ** 2)
        ks = np.max(np.abs(np.arange(1, len(c2) + 1) / len(c2) - c2))
        assert ks < 0.01, ks
        v = _sphere(20000, 5)
        v = v[v[:, 2] > 0.2]
        f = api.ms_diffuse_eval(wi, v, 0.02, rho, dist=dist, seed=3).astype(np.float64)
        want = rho[None, :] * v[:, 2:3] / math.pi
        # one event about a visible normal within ~alpha of the mean normal:
        # each estimate scatters by ~alpha tan(theta_o), the mean has no bias
        r = f / want
        assert abs(r.mean() - 1.0) < 3e-3 and r.std() < 0.15, (r.mean(), r.std())


@pytest.mark.parametrize("dist", [1, 0])
def test_diffuse_sampler_structure(api, dist):
    rho = np.array([0.9, 0.6, 0.3])
    wo, wm, w, order = api.ms_diffuse_sample(_wi(60.0), 0.8, rho, 50000, dist=dist, seed=11)
    np.testing.assert_allclose(w, rho[None, :] ** order[:, None], rtol=2e-6 * 100, atol=1e-30)
    assert (order > 1).mean() > 0.05
    assert np.all(np.sum(wo * wm, axis=1) > -1e-5)
    np.testing.assert_allclose(np.linalg.norm(wm, axis=1), 1.0, atol=2e-6)
