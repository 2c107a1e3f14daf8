"""Pins the oracle's microsurface random walk (Heitz et al. 2016, restated from
ORF/MicrofacetGGX.h:254-305, 412-438) against results of that paper that do
not share its code path:

- a dielectric walk has no absorption: every walk ends (within the 10-event
  cap, all but a tiny fraction) with a unit exit direction, on the incident
  side after reflection or on the far side after transmission;
- its single-scattering part is the microfacet BSDF with the height-correlated
  Smith masking-shadowing G2 = 1 / (1 + Lambda(wi) + Lambda(wo)): the density
  of exits after exactly one reflection is F(wi.h) D(h) G2 / (4 cos_i).  That
  density is evaluated here in numpy (GGX D and Lambda, Fresnel equations) and
  compared by binned total variation; the separable G1(wi) G1(wo) (a plausible
  mistake in the height sampling) moves the mass by far more than the noise.
"""
import math

import numpy as np
import pytest

GL = np.polynomial.legendre.leggauss


def _fresnel(c, eta):
    st2 = eta * eta * (1.0 - c * c)
    ct = np.sqrt(np.maximum(0.0, 1.0 - st2))
    rs = (eta * c - ct) / (eta * c + ct)
    rp = (c - eta * ct) / (c + eta * ct)
    return np.where(st2 >= 1.0, 1.0, 0.5 * (rs * rs + rp * rp))


def _lam(cos_t, a):
    t2 = (1.0 - cos_t**2) / cos_t**2
    return 0.5 * (-1.0 + np.sqrt(1.0 + a * a * t2))


def _D(cm, a):
    t2 = (1.0 - cm**2) / cm**2
    return 1.0 / (math.pi * a * a * cm**4 * (1.0 + t2 / (a * a)) ** 2)


def _single_refl_density(wi, wo, a, eta, separable=False):
    h = wi[None, :] + wo
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    F = _fresnel(np.abs(wo @ np.zeros(3) + (h @ wi)), eta)
    li, lo = _lam(wi[2], a), _lam(wo[:, 2], a)
    G2 = 1.0 / (1.0 + li) / (1.0 + lo) if separable else 1.0 / (1.0 + li + lo)
    return F * _D(h[:, 2], a) * G2 / (4.0 * wi[2])


def _bin_masses(wi, a, eta, cb, pb, separable=False, q=8):
    xg, wg = GL(q)
    m = np.zeros((len(cb) - 1, len(pb) - 1))
    for i in range(len(cb) - 1):
        c = 0.5 * (cb[i + 1] - cb[i]) * xg + 0.5 * (cb[i + 1] + cb[i])
        wc = 0.5 * (cb[i + 1] - cb[i]) * wg
        for j in range(len(pb) - 1):
            p = 0.5 * (pb[j + 1] - pb[j]) * xg + 0.5 * (pb[j + 1] + pb[j])
            wp = 0.5 * (pb[j + 1] - pb[j]) * wg
            C, P = np.meshgrid(c, p, indexing="ij")
            s = np.sqrt(1.0 - C**2)
            wo = np.stack([s * np.cos(P), s * np.sin(P), C], -1).reshape(-1, 3)
            m[i, j] = (_single_refl_density(wi, wo, a, eta, separable) * np.outer(wc, wp).ravel()).sum()
    return m


@pytest.mark.parametrize("alpha,theta,phi", [(0.6, 1.3, 0.3), (0.9, 1.1, 2.0), (0.3, 0.8, 0.3)])
def test_walk_single_scattering_is_height_correlated_microfacet_brdf(orc, alpha, theta, phi):
    eta = 1.0 / 1.5
    wi = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])
    n = 1 << 22
    wo, order, ok = orc.ggx_ms_sample(wi.astype(np.float32), eta, alpha, n, seed=7)
    wo = wo.astype(np.float64)
    assert ok.mean() > 0.999
    assert np.abs(np.linalg.norm(wo[ok == 1], axis=1) - 1.0).max() < 1e-5
    sel = (order == 1) & (wo[:, 2] > 0) & (ok == 1)
    cb = np.linspace(0.0, 1.0, 17) ** 0.5
    pb = np.linspace(-math.pi, math.pi, 17)
    cnt, _, _ = np.histogram2d(wo[sel, 2], np.arctan2(wo[sel, 1], wo[sel, 0]), bins=[cb, pb])
    want = _bin_masses(wi, alpha, eta, cb, pb)
    want_sep = _bin_masses(wi, alpha, eta, cb, pb, separable=True)
    # shape: binned total variation at the noise level (~5e-4 at 2^22 walks)
    tv = 0.5 * np.abs(cnt / n - want).sum()
    assert tv < 0.002, tv
    # mass: the single-reflection albedo within 4 sigma of the height-correlated
    # value; the separable G2 is many sigma away at grazing incidence
    mass, sigma = sel.mean(), math.sqrt(sel.mean() / n)
    assert abs(mass - want.sum()) < 4 * sigma + 1e-5, (mass, want.sum(), sigma)
    if theta > 1.0:
        assert abs(mass - want_sep.sum()) > 15 * sigma, (mass, want_sep.sum(), sigma)


def test_walk_exits_on_the_side_its_last_event_chose(orc):
    wi = np.array([0.3, 0.2, math.sqrt(1 - 0.13)], np.float32)
    wo, order, ok = orc.ggx_ms_sample(wi, 1.0 / 1.5, 0.6, 1 << 18, seed=3)
    assert (order >= 1).all() and ok.mean() > 0.999
    # reflection exits go up, transmission exits go down; both happen
    w = wo[ok == 1]
    up, down = (w[:, 2] > 0).mean(), (w[:, 2] < 0).mean()
    assert up > 0.02 and down > 0.5 and abs(up + down - 1.0) < 1e-6
    # a walk started from inside exits too (the walk is symmetric in the sides)
    wo2, order2, ok2 = orc.ggx_ms_sample(-wi, 1.0 / 1.5, 0.6, 1 << 16, seed=4)
    # from the denser side total internal reflection lengthens walks: 0.2% reach the cap
    assert ok2.mean() > 0.995 and (order2 >= 1).all()
