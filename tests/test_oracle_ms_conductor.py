"""Pins the oracle's GGX conductor microsurface walk
(ggx_multiscattering_conductor_BSDF_sample, ORF/MicrofacetGGX.h:614-637), CPU only.

- A perfect mirror microsurface (complex ior 1 + 1e4 i: Fresnel 1 to 1e-7 at
  every angle) loses no energy: every completed walk leaves upwards with
  weight 1, and the walk completes within 10 events almost always.
- At low roughness the walk is single scattering: 99% of walks end after one
  event at alpha 0.05.
- Both the GGX and the Beckmann walks (ORF/MicrofacetBeckmann.h:711-736).
- The Beckmann dielectric walk at roughness 0.02 reflects the Fresnel share.
- Energy lost by a real conductor grows with the number of bounces: the mean
  weight of order-k walks is below that of order-(k-1) walks for gold-like ior.
- The Beckmann conductor walk's single-scattering part: the reference's
  Beckmann walk leaves the microsurface after an event with the
  height-averaged probability 1 / (1 + Lambda(wo)) (ORF/MicrofacetBeckmann.h:
  284-306), so a mirror walk's exits after exactly one event follow the
  visible-normal reflection times that escape probability, i.e. the Beckmann
  BRDF with the separable Smith term G1(wi) G1(wo):
  D(h) G1(wi) G1(wo) / (4 cos_i), with Lambda(a) = (erf(a) - 1) / 2 +
  exp(-a^2) / (2 a sqrt(pi)), a = 1 / (alpha tan theta), and the Beckmann NDF
  in closed form.  Compared by binned total variation and by the order-1
  mass; the height-correlated G2 = 1 / (1 + Lambda(wi) + Lambda(wo)) is many
  sigma away, and a wrong Lambda (e.g. GGX's) further still.
"""
import math

import numpy as np
import pytest

from tests.oracle_api import Oracle

MIRROR = ((1.0, 1e4),) * 3
GOLD = ((0.18, 3.42), (0.42, 2.35), (1.37, 1.77))


def _wi(theta_deg):
    t = np.radians(theta_deg)
    return np.array([np.sin(t), 0.0, np.cos(t)], np.float32)


@pytest.mark.parametrize("dist", [1, 0])
def test_mirror_walk_conserves_energy(dist):
    o = Oracle()
    for theta, alpha in ((10, 0.3), (60, 0.6), (80, 0.9)):
        wo, w, order, ok = o.ggx_ms_conductor(_wi(theta), MIRROR, alpha, 20000, dist=dist)
        done = ok == 1
        assert done.mean() > 0.999, (theta, alpha, done.mean())
        assert np.all(wo[done, 2] > 0.0)
        np.testing.assert_allclose(w[done], 1.0, atol=1e-5)
        assert np.all(order[done] >= 1)
        if alpha >= 0.6:
            assert (order[done] > 1).mean() > 0.02  # multiple scattering happens at high roughness


@pytest.mark.parametrize("dist", [1, 0])
def test_low_roughness_is_single_scattering(dist):
    o = Oracle()
    _, _, order, ok = o.ggx_ms_conductor(_wi(30), MIRROR, 0.05, 20000, dist=dist)
    assert (order[ok == 1] == 1).mean() > 0.99


@pytest.mark.parametrize("dist", [1, 0])
def test_gold_loses_energy_per_bounce(dist):
    o = Oracle()
    wo, w, order, ok = o.ggx_ms_conductor(_wi(70), GOLD, 0.8, 40000, dist=dist)
    done = ok == 1
    m1 = w[done & (order == 1)].mean(axis=0)
    m2 = w[done & (order == 2)].mean(axis=0)
    m3 = w[done & (order == 3)].mean(axis=0)
    assert np.all(m1 > m2) and np.all(m2 > m3), (m1, m2, m3)
    assert np.all((w[done] >= 0) & (w[done] <= 1))


def test_beckmann_dielectric_walk_smooth_limit_is_fresnel():
    """Beckmann dielectric walk (ORF/MicrofacetBeckmann.h:445-475) at roughness
    0.02: the share of walks leaving upwards is the unpolarised Fresnel
    reflectance at the incidence angle (45 degrees, relative index 1/1.5)."""
    o = Oracle()
    wo, order, ok = o.beckmann_ms_sample(_wi(45), 1.0 / 1.5, 0.02, 40000)
    up = (wo[ok == 1, 2] > 0).mean()
    ci = math.cos(math.radians(45))
    st = math.sin(math.radians(45)) / 1.5
    ct = math.sqrt(1 - st * st)
    rs = (ci - 1.5 * ct) / (ci + 1.5 * ct)
    rp = (1.5 * ci - ct) / (1.5 * ci + ct)
    want = 0.5 * (rs * rs + rp * rp)
    assert abs(up - want) < 0.05 * want, (up, want)
    assert (ok == 1).mean() > 0.999


def _beckmann_lambda(cos_t, a):
    from scipy.special import erf

    tan_t = np.sqrt(np.maximum(1.0 - cos_t**2, 1e-300)) / cos_t
    x = 1.0 / (a * tan_t)
    return (erf(x) - 1.0) / 2.0 + np.exp(-x * x) / (2.0 * x * math.sqrt(math.pi))


def _beckmann_D(cm, a):
    t2 = (1.0 - cm**2) / cm**2
    return np.exp(-t2 / (a * a)) / (math.pi * a * a * cm**4)


def _beckmann_order1_masses(wi, a, cb, pb, separable=False, q=8):
    xg, wg = np.polynomial.legendre.leggauss(q)
    li = _beckmann_lambda(wi[2], a)
    m = np.zeros((len(cb) - 1, len(pb) - 1))
    for i in range(len(cb) - 1):
        c = 0.5 * (cb[i + 1] - cb[i]) * xg + 0.5 * (cb[i + 1] + cb[i])
        wc = 0.5 * (cb[i + 1] - cb[i]) * wg
        for j in range(len(pb) - 1):
            p = 0.5 * (pb[j + 1] - pb[j]) * xg + 0.5 * (pb[j + 1] + pb[j])
            wp = 0.5 * (pb[j + 1] - pb[j]) * wg
            C, P = np.meshgrid(c, p, indexing="ij")
            sn = np.sqrt(1.0 - C**2)
            wo = np.stack([sn * np.cos(P), sn * np.sin(P), C], -1).reshape(-1, 3)
            h = wi[None, :] + wo
            h /= np.linalg.norm(h, axis=1, keepdims=True)
            lo = _beckmann_lambda(wo[:, 2], a)
            G2 = 1.0 / ((1.0 + li) * (1.0 + lo)) if separable else 1.0 / (1.0 + li + lo)
            dens = _beckmann_D(h[:, 2], a) * G2 / (4.0 * wi[2])
            m[i, j] = (dens * np.outer(wc, wp).ravel()).sum()
    return m


@pytest.mark.parametrize("alpha,theta", [(0.5, 70.0), (0.8, 75.0), (0.4, 80.0)])
def test_beckmann_walk_single_scattering_is_separable_smith_brdf(alpha, theta):
    o = Oracle()
    wi = _wi(theta).astype(np.float64)
    n = 1 << 22
    wo, w, order, ok = o.ggx_ms_conductor(wi.astype(np.float32), MIRROR, alpha, n, seed=13, dist=0)
    wo = wo.astype(np.float64)
    sel = (order == 1) & (ok == 1) & (wo[:, 2] > 0)
    cb = np.linspace(0.0, 1.0, 13) ** 0.5
    pb = np.linspace(-math.pi, math.pi, 13)
    cnt, _, _ = np.histogram2d(wo[sel, 2], np.arctan2(wo[sel, 1], wo[sel, 0]), bins=[cb, pb])
    want = _beckmann_order1_masses(wi, alpha, cb, pb, separable=True)
    tv = 0.5 * np.abs(cnt / n - want).sum()
    assert tv < 0.003, tv
    mass, sigma = sel.mean(), math.sqrt(sel.mean() * (1 - sel.mean()) / n)
    assert abs(mass - want.sum()) < 4 * sigma + 1e-4, (mass, want.sum(), sigma)
    hc = _beckmann_order1_masses(wi, alpha, cb, pb).sum()
    assert abs(mass - hc) > 6 * sigma, (mass, hc, sigma)
