"""Pins the oracle's smooth glass (CU/transparent_shader.cu:51-114) and normal
display (CU/normal_shader.cu) materials, CPU only.

- Glass: a camera looking at 45 degrees onto a one-sided glass plane (index
  2.5) under an environment that is 1 above the horizon and 0 below sees the
  reflected fraction only, so the pixel mean is the unpolarised Fresnel
  reflectance (Rs + Rp) / 2 of the textbook equations at 45 degrees.
- Glass in a white furnace: an absorption-free glass sphere under a constant
  environment of 1 renders 1 (every path leaves with weight 1).
- Beer absorption: light crossing a glass slab of thickness L at 45 degrees
  towards an environment that is bright only below the horizon carries
  sum_k T^2 R^2k exp(-sigma L (2k + 1) / cos theta_t) (Russian roulette on the
  mean keeps it unbiased).
- Normal display: a plane facing +y shows (0.5, 1, 0.5).
- Shadows pass through normal-display objects (their any-hit does not clear the
  attenuation): a ground point under one renders (in green, which the
  blocker's downward normal colour lacks) what it renders without it, and 0
  when the blocker is diffuse.
"""
import math

import numpy as np
import pytest

from open_rendering_framework_amd import scenes
from tests.oracle_scene import OracleScene


def _half_env(w=128, h=64):
    env = np.zeros((h, w, 4), np.float32)
    env[h // 2:, :, :3] = 1.0  # rows h/2..h-1: theta < 90 deg (above the horizon)
    env[..., 3] = 1.0
    return env


def _fresnel_unpolarised(ci, n):
    st = math.sqrt(1 - ci * ci) / n
    ct = math.sqrt(1 - st * st)
    rs = (ci - n * ct) / (ci + n * ct)
    rp = (n * ci - ct) / (n * ci + ct)
    return 0.5 * (rs * rs + rp * rp)


def _plane(y, half, mat):
    v, t = scenes.quad(y, half, facing_down=False)
    return scenes.Mesh(v, t, mat, normals=np.tile([[0, 1, 0]], (4, 1)).astype(np.float32))


def test_glass_reflectance_is_fresnel():
    glass = scenes.Material("transparent_shader.cu", floats={"ior": (2.5,)})
    d = scenes.SceneDesc([_plane(0.0, 10.0, 0)], [glass], [], _half_env(), False)
    d.eye, d.U, d.V, d.W = scenes.pinhole((0.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.01, 9, 9)
    d.max_depth = 0
    img, _ = OracleScene(d, fp64=True).render(9, 9, 0, 2000)
    got = (img[..., 0] / img[..., 3]).mean()
    want = _fresnel_unpolarised(math.cos(math.pi / 4), 2.5)
    assert abs(got - want) < 0.02 * want, (got, want)


def test_glass_sphere_white_furnace():
    sv, sn, st = scenes.icosphere(3, amplitude=0.0)
    glass = scenes.Material("transparent_shader.cu", floats={"ior": (1.5,)})
    env = np.ones((32, 64, 4), np.float32)
    d = scenes.SceneDesc([scenes.Mesh(sv, st, 0, normals=sn)], [glass], [], env, False)
    d.eye, d.U, d.V, d.W = scenes.pinhole((0, 0, 3.0), (0, 0, 0), (0, 1, 0), 12.0, 16, 16)
    d.max_depth = 64
    img, _ = OracleScene(d, fp64=True).render(16, 16, 0, 64)
    v = img[..., :3] / img[..., 3:4]
    assert abs(v.mean() - 1.0) < 2e-3, v.mean()


def test_glass_slab_beer_absorption():
    sigma, L, n = 0.7, 0.5, 1.5
    top, bottom = _plane(0.0, 10.0, 0), _plane(-L, 10.0, 0)
    bottom.tris = bottom.tris[:, ::-1].copy()  # faces -y: outward normal of the slab's lower face
    bottom.normals = -bottom.normals
    glass = scenes.Material("transparent_shader.cu", floats={"ior": (n,), "glass_absorption": (sigma,) * 3})
    env = np.zeros((64, 128, 4), np.float32)
    env[: 32, :, :3] = 1.0  # below the horizon: the light the ray meets after the slab
    d = scenes.SceneDesc([top, bottom], [glass], [], env, False)
    # 45 degree incidence (a straight-down exit would sample the environment's pole)
    d.eye, d.U, d.V, d.W = scenes.pinhole((0.0, 2.0, 2.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.01, 9, 9)
    d.max_depth = 64
    img, _ = OracleScene(d, fp64=True).render(9, 9, 0, 4000)
    got = (img[..., 0] / img[..., 3]).mean()
    ci = math.cos(math.pi / 4)
    ct = math.sqrt(1 - (1 - ci * ci) / (n * n))
    R = _fresnel_unpolarised(ci, n)  # the same at the inner faces (reciprocity)
    T = 1.0 - R
    # the paths leaving through the lower face after 2k internal reflections,
    # each crossing absorbing exp(-sigma L / cos theta_t)
    want = sum(T * T * (R * R) ** k * math.exp(-sigma * L * (2 * k + 1) / ct) for k in range(20))
    assert abs(got - want) < 0.02 * want, (got, want)


def test_normal_display_colour():
    nm = scenes.Material("normal_shader.cu")
    d = scenes.SceneDesc([_plane(0.0, 5.0, 0)], [nm], [], None, False)
    d.eye, d.U, d.V, d.W = scenes.pinhole((0.0, 2.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 20.0, 4, 4)
    img, _ = OracleScene(d).render(4, 4, 0, 2)
    np.testing.assert_array_equal(img[..., :3] / img[..., 3:4], np.broadcast_to([0.5, 1.0, 0.5], (4, 4, 3)))


@pytest.mark.parametrize("blocker,expect", [("normal_shader.cu", "same"), ("diffuse_shader.cu", "dark")])
def test_shadow_rays_pass_through_normal_display(blocker, expect):
    ground = _plane(-1.0, 10.0, 0)
    bv, bt = scenes.quad(1.0, 0.5, facing_down=True)
    block = scenes.Mesh(bv, bt, 1, normals=np.tile([[0, -1, 0]], (4, 1)).astype(np.float32))
    mats = [scenes.Material("diffuse_shader.cu", floats={"diffuse_color": (0.5, 0.5, 0.5)}),
            scenes.Material(blocker)]
    light = [scenes.Light("point", (10.0, 10.0, 10.0), position=(0.0, 3.0, 0.0))]
    cam = scenes.pinhole((2.0, 0.5, 2.0), (0.0, -1.0, 0.0), (0.0, 1.0, 0.0), 0.5, 4, 4)
    with_b = scenes.SceneDesc([ground, block], mats, light, None, False, *cam, 4, 4, 0)
    without = scenes.SceneDesc([ground], mats, light, None, False, *cam, 4, 4, 0)
    a, _ = OracleScene(with_b).render(4, 4, 0, 4)
    b, _ = OracleScene(without).render(4, 4, 0, 4)
    assert b[..., :3].min() > 0
    if expect == "same":
        # green only: a bounce that reaches the blocker's underside sees its
        # normal colour (0.5, 0, 0.5), which has no green
        np.testing.assert_array_equal(a[..., 1], b[..., 1])
    else:
        assert a[..., :3].max() == 0


def _flat_scene(floats=None, uints=None, blocker=False):
    """A flat ground under a point light straight above; optionally a diffuse
    blocker between them.  The camera looks at the ground below the light."""
    ground = _plane(-1.0, 10.0, 0)
    meshes = [ground]
    mats = [scenes.Material("flat_shader.cu", floats=floats or {}, uints=uints or {}),
            scenes.Material("diffuse_shader.cu")]
    if blocker:
        bv, bt = scenes.quad(1.0, 0.5, facing_down=True)
        meshes.append(scenes.Mesh(bv, bt, 1))
    light = [scenes.Light("point", (10.0, 10.0, 10.0), position=(0.0, 3.0, 0.0))]
    cam = scenes.pinhole((1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (0.0, 1.0, 0.0), 0.5, 4, 4)
    return scenes.SceneDesc(meshes, mats, light, None, False, *cam, 4, 4, 0)


@pytest.mark.parametrize("shadow,blocker,want", [(0, True, (0.2, 0.4, 0.6)), (1, False, (0.2, 0.4, 0.6)),
                                                 (1, True, (0.0, 0.0, 0.0))])
def test_flat_colour_and_shadow(shadow, blocker, want):
    d = _flat_scene({"diffuse_color": (0.2, 0.4, 0.6)}, {"flat_shadow": (shadow,)}, blocker)
    img, _ = OracleScene(d).render(4, 4, 0, 2)
    np.testing.assert_allclose(img[..., :3] / img[..., 3:4], np.broadcast_to(want, (4, 4, 3)), rtol=1e-6)


@pytest.mark.parametrize("threshold,want", [(0.0, (1.0, 0.5, 0.0)), (1e30, (0.2, 0.4, 0.6))])
def test_flat_highlight_threshold(threshold, want):
    """A visible light with a positive specular term shows the highlight colour
    at threshold 0 and never at an unreachable threshold."""
    d = _flat_scene({"diffuse_color": (0.2, 0.4, 0.6), "highlight_color": (1.0, 0.5, 0.0),
                     "highlight_threshold": (threshold,), "roughness": (0.5, 0.5)}, {"highlight": (1,)})
    img, _ = OracleScene(d).render(4, 4, 0, 2)
    np.testing.assert_allclose(img[..., :3] / img[..., 3:4], np.broadcast_to(want, (4, 4, 3)), rtol=1e-6)
