"""Pins the rough translucent boundary's NEE reflection term (CPU only).

CU/rough_subsurface_scattering_shader.cu:146-165 weights a light sample seen
from the outside by c(i, l) = F(l.h) G1(i, h) G1(l, h) D_wi(h) / (4 |i.n| |l.n|)
with h the half vector of i and l and D_wi the visible-normal density (the
reference's choice, kept).  Two independent routes fix the oracle's term:

- the integral over light directions, taken by quadrature over l on a fine
  grid of the oracle's own values, equals the integral over half vectors
  int F G1(i,h) G1(l,h) D_wi(h) (i.h) / (|i.n| |l.n|) dw_h (the Jacobian of
  reflection, dw_l = 4 (l.h) dw_h), taken by quadrature over h with the GGX
  NDF and Smith G1 in closed form (Heitz 2014) and the Fresnel equations
  written here; a dropped or squared factor, D in place of D_wi, the wrong
  cosine in the denominator or a wrong Fresnel angle breaks the equality;
- as the roughness goes to 0 the lobe becomes the mirror direction and the
  integral tends to F(i.n) / (i.n): the smooth boundary's Fresnel reflectance
  over the cosine the reference's estimator leaves out.
"""
import math

import numpy as np
import pytest

from tests.oracle_scene import OracleScene  # noqa: F401  (loads the oracle library)
from tests import oracle_scene

ETA = 1.0 / 1.3  # n1 / n2 as the boundary is entered from outside


def _lib():
    import ctypes

    L = oracle_scene.oracle.lib(fp64=True)
    L.orc_rough_nee_reflect_batch.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                              ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_void_p]
    return L


def _oracle_c(L, wi, wl, ax, ay, dist):
    wl = np.ascontiguousarray(wl, np.float32)
    wis = np.ascontiguousarray(np.broadcast_to(np.asarray(wi, np.float32), wl.shape))
    out = np.zeros(len(wl))
    L.orc_rough_nee_reflect_batch(len(wl), wis.ctypes.data, wl.ctypes.data, ax, ay, dist, ETA, out.ctypes.data)
    return out


def _fresnel(c, eta):
    st2 = eta * eta * (1.0 - c * c)
    ct = np.sqrt(np.maximum(0.0, 1.0 - st2))
    rs = (eta * c - ct) / (eta * c + ct)
    rp = (c - eta * ct) / (c + eta * ct)
    return np.where(st2 >= 1.0, 1.0, 0.5 * (rs * rs + rp * rp))


def _ggx_D(h, ax, ay):
    t2 = (h[..., 0] ** 2 / ax ** 2 + h[..., 1] ** 2 / ay ** 2) / h[..., 2] ** 2
    return 1.0 / (math.pi * ax * ay * h[..., 2] ** 4 * (1.0 + t2) ** 2)


def _ggx_G1(v, h, ax, ay):
    t2 = (v[..., 0] ** 2 * ax ** 2 + v[..., 1] ** 2 * ay ** 2) / v[..., 2] ** 2
    g = 2.0 / (1.0 + np.sqrt(1.0 + t2))
    return np.where((np.sum(v * h, -1) / v[..., 2]) > 0, g, 0.0)


def _grid(n_t, n_p, tmax=math.pi / 2):
    t = (np.arange(n_t) + 0.5) * tmax / n_t
    p = (np.arange(n_p) + 0.5) * 2 * math.pi / n_p
    T, P = np.meshgrid(t, p, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (tmax / n_t) * (2 * math.pi / n_p)).ravel()
    return d, dw


def _integral_over_h(wi, ax, ay):
    h, dw = _grid(1500, 1200)
    ih = h @ wi
    l = 2.0 * ih[:, None] * h - wi
    ok = (l[:, 2] > 0) & (ih > 0)
    h, l, ih, dw = h[ok], l[ok], ih[ok], dw[ok]
    gi = _ggx_G1(np.broadcast_to(wi, h.shape), h, ax, ay)
    dwi = gi * ih * _ggx_D(h, ax, ay) / wi[2]
    f = _fresnel(ih, ETA) * gi * _ggx_G1(l, h, ax, ay) * dwi * ih / (wi[2] * l[:, 2])
    return float((f * dw).sum())


@pytest.mark.parametrize("ax,ay,theta", [(0.3, 0.3, 20.0), (0.3, 0.3, 60.0), (0.6, 0.6, 45.0), (0.2, 0.5, 35.0)])
def test_nee_reflection_change_of_variables(ax, ay, theta):
    L = _lib()
    t = math.radians(theta)
    wi = np.array([math.sin(t) * math.cos(0.4), math.sin(t) * math.sin(0.4), math.cos(t)])
    wi32 = wi.astype(np.float32).astype(np.float64)
    wl, dw = _grid(1500, 1200)
    c = _oracle_c(L, wi32, wl, ax, ay, 1)
    over_l = float((c * dw).sum())
    over_h = _integral_over_h(wi32, ax, ay)
    assert abs(over_l - over_h) < 2e-3 * over_h, (over_l, over_h)


@pytest.mark.parametrize("theta", [15.0, 50.0, 70.0])
def test_nee_reflection_smooth_limit(theta):
    L = _lib()
    t = math.radians(theta)
    wi = np.array([math.sin(t), 0.0, math.cos(t)])
    # a fine grid around the mirror direction holds the lobe at alpha 0.002 up to
    # its GGX tail beyond the window, ~(alpha / 0.04)^2 = 0.25%
    a = 0.002
    mirror = np.array([-wi[0], 0.0, wi[2]])
    k = np.linspace(-0.08, 0.08, 801)
    du = k[1] - k[0]
    U, V = np.meshgrid(k, k, indexing="ij")
    e1 = np.cross(mirror, [0.0, 1.0, 0.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(mirror, e1)
    d = mirror[None, :] + U.reshape(-1, 1) * e1 + V.reshape(-1, 1) * e2
    nrm = np.linalg.norm(d, axis=1)
    wl = d / nrm[:, None]
    dw = du * du / nrm ** 3  # solid angle of a tangent-plane cell
    c = _oracle_c(L, wi, wl, a, a, 1)
    got = float((c * dw).sum())
    want = float(_fresnel(np.array(wi[2]), ETA)) / wi[2]
    assert abs(got - want) < 0.01 * want, (got, want)
