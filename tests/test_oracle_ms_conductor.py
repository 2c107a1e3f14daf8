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


def test_low_roughness_is_single_scattering():
    o = Oracle()
    _, _, order, ok = o.ggx_ms_conductor(_wi(30), MIRROR, 0.05, 20000)
    assert (order[ok == 1] == 1).mean() > 0.99


def test_gold_loses_energy_per_bounce():
    o = Oracle()
    wo, w, order, ok = o.ggx_ms_conductor(_wi(70), GOLD, 0.8, 40000)
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
