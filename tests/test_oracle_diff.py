"""Oracle of the differentiable form (textured albedo / roughness, detached sampling).

Pins:
  * constant textures equal to the material values reproduce the plain render
    bit for bit (the texel substitution changes nothing else);
  * re-weighting: with sampling and Russian roulette frozen at theta0 and the
    BSDF evaluated at theta, the estimator's mean is the plain render's mean at
    theta (importance-sampling identity) -- checked for the albedo (RR prob from
    theta0) and the GGX roughness (VNDF drawn at alpha0, weight
    F G1(wi) G1(wo) D / (G1_0(wi) D_0)); a dropped G1 or D term fails it;
  * the loss is linear in the image gradient (sum of two losses == loss of sum);
  * the texel lookup (no reference counterpart): points at the texel centres of
    the planar rectangle and of the sphere's (theta, phi) grid land on their
    texels, points outside clamp to the edge texels (a swapped u/v, a wrong
    atan2 argument order or an off-by-one fails it).
"""
import numpy as np
import pytest

from open_rendering_framework_amd import scenes
from tests.oracle_scene import OracleScene

W = 16


def _scene(**kw):
    return scenes.path_scene(width=W, height=W, level=3, env_size=(64, 32), **kw)


def _views(spp, n=2, frame0=0):
    return scenes.ring_views(n=n, size=W, spp=spp, frame0=frame0)


def test_constant_textures_reproduce_plain_render():
    d = _scene()
    orc = OracleScene(d)
    views = _views(4, n=1)
    v = views[0]
    # the plain render with the view's camera
    d.eye, d.U, d.V, d.W = v["eye"], v["U"], v["V"], v["W"]
    plain = OracleScene(d)
    film, _ = plain.render(W, W, v["frame0"], v["spp"])
    want = film[..., :3] / np.maximum(film[..., 3:], 1)
    alb = np.full((4, 4, 3), 0.5)
    rough = np.full((4, 4), 0.3)
    orc.diff_set((4, 4), scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, alb.astype(np.float32),
                 rough.astype(np.float32))
    _, img = orc.diff_eval(views)
    np.testing.assert_array_equal(img[0], want)


def _mean_and_se(orc, views, batches):
    """Image-average radiance over `batches` disjoint frame ranges: (mean, std error)."""
    vals = []
    for b in range(batches):
        _, img = orc.diff_eval(views, shift=b * 1000)
        vals.append(img.mean())
    vals = np.array(vals)
    return vals.mean(), vals.std(ddof=1) / np.sqrt(len(vals))


@pytest.mark.parametrize("which", ["albedo", "roughness"])
def test_frozen_sampling_reweighting_is_unbiased(which):
    tex = 4
    views = _views(32)
    if which == "albedo":
        a0, a1 = 0.5, 0.7
        d_plain = _scene()
        d_plain.materials[0].floats["diffuse_color"] = (a1, a1, a1)
        frozen = np.full((tex, tex, 3), a0)
        evalv = np.full((tex, tex, 3), a1)
        kw = dict(albedo=evalv, albedo_frozen=frozen, roughness=np.full((tex, tex), 0.3))
    else:
        a0, a1 = 0.3, 0.7
        d_plain = _scene(alpha=a1)
        frozen = np.full((tex, tex), a0)
        evalv = np.full((tex, tex), a1)
        kw = dict(albedo=np.full((tex, tex, 3), 0.5), roughness=evalv, roughness_frozen=frozen)
    rew = OracleScene(_scene())
    rew.diff_set((tex, tex), scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, **kw)
    plain = OracleScene(d_plain)
    plain.diff_set((tex, tex), -1, -1)
    m_rew, se_rew = _mean_and_se(rew, views, 16)
    m_pl, se_pl = _mean_and_se(plain, views, 16)
    se = np.hypot(se_rew, se_pl)
    assert abs(m_rew - m_pl) < 4.0 * se + 1e-6, (m_rew, m_pl, se)
    # and the re-weighted estimator is not the theta0 render (the check has power)
    base = OracleScene(_scene())
    base.diff_set((tex, tex), -1, -1)
    m0, se0 = _mean_and_se(base, views, 16)
    assert abs(m0 - m_pl) > 4.0 * np.hypot(se0, se_pl), (m0, m_pl)


def test_loss_is_linear_in_image_gradient():
    orc = OracleScene(_scene())
    a, r = scenes.diff_textures(4, seed=3)
    orc.diff_set((4, 4), scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, a, r)
    views = _views(2)
    rng = np.random.default_rng(1)
    g1 = rng.standard_normal((2, W, W, 3))
    g2 = rng.standard_normal((2, W, W, 3))
    l1, _ = orc.diff_eval(views, shift=7, dimg=g1, images=False)
    l2, _ = orc.diff_eval(views, shift=7, dimg=g2, images=False)
    l12, _ = orc.diff_eval(views, shift=7, dimg=g1 + g2, images=False)
    assert abs(l12 - (l1 + l2)) <= 1e-9 * (abs(l1) + abs(l2)) + 1e-9
    _, img = orc.diff_eval(views, shift=7)
    # loss == sum(dimg * image) when every sample is finite
    np.testing.assert_allclose(l1, (g1 * img).sum(), rtol=1e-9)


SSS0 = np.array([0.0011, 0.0024, 0.014, 2.55, 3.21, 3.77])
MILK = 1


def _milk_scene(theta):
    d = scenes.subsurface_scene(width=W, height=W, level=3, glass_level=1, env_size=(64, 32), probes=0)
    d.materials[MILK].floats["absorption"] = tuple(theta[:3])
    d.materials[MILK].floats["scattering"] = tuple(theta[3:])
    return d


def _milk_oracle(theta, probes):
    o = OracleScene(_milk_scene(theta))
    o.probes_set(probes, 4, MILK)
    o.probes_values(probes["L"], probes["w12"], probes["sn"])
    return o


def test_scattering_reweighting_is_unbiased():
    """The translucent material's coefficients as parameters: shaded at theta1,
    with the gather's exit points and the extinction roulette drawn at theta0,
    the estimator's mean is the plain render's at theta1 (same probe values on
    both sides, the probes being frozen); and differs from the theta0 render."""
    theta1 = SSS0 * np.array([8.0, 8.0, 8.0, 0.6, 0.6, 0.6])
    m = _milk_scene(SSS0).meshes[MILK]
    n = len(m.verts)
    area = 0.5 * np.linalg.norm(np.cross(m.verts[m.tris[:, 1]] - m.verts[m.tris[:, 0]],
                                         m.verts[m.tris[:, 2]] - m.verts[m.tris[:, 0]]), axis=1).sum()
    pad = lambda a: np.concatenate([a, np.zeros((len(a), 1), np.float32)], 1).astype(np.float32)  # noqa: E731
    probes = {"pos": pad(m.verts), "nrm": pad(m.normals), "cell": float(2.5 * np.sqrt(area / n))}
    src = OracleScene(_milk_scene(SSS0))
    src.probes_set(probes, 4, MILK)
    L, w12, sn = src.probe_pass(0, normals=True)
    probes.update(L=L, w12=w12, sn=sn)
    views = scenes.ring_views(n=1, size=W, spp=16, radius=2.5, height=0.8)
    rew = _milk_oracle(SSS0, probes)
    rew.diff_set_sss(MILK, theta1, SSS0)
    plain1 = _milk_oracle(theta1, probes)
    plain0 = _milk_oracle(SSS0, probes)
    m_rew, se_rew = _mean_and_se(rew, views, 16)
    m_1, se_1 = _mean_and_se(plain1, views, 16)
    m_0, se_0 = _mean_and_se(plain0, views, 16)
    assert abs(m_rew - m_1) < 4.0 * np.hypot(se_rew, se_1) + 1e-7, (m_rew, m_1, se_rew, se_1)
    assert abs(m_0 - m_1) > 4.0 * np.hypot(se_0, se_1), (m_0, m_1)


def _texel_lookup(mode, uvp, pts, tw, th, fp64):
    import ctypes

    import oracle

    L = oracle.lib(fp64=fp64)
    L.orc_diff_texel_batch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_void_p]
    uvp = np.ascontiguousarray(uvp, np.float64)
    pts = np.ascontiguousarray(pts, np.float64)
    out = np.zeros(len(pts), np.int32)
    L.orc_diff_texel_batch(len(pts), mode, uvp.ctypes.data, pts.ctypes.data, tw, th, out.ctypes.data)
    return out


@pytest.mark.parametrize("fp64", [False, True])
def test_planar_texel_lookup_hits_texel_centres(fp64):
    tw, th = 7, 5
    x0, z0, x1, z1 = -1.5, 0.25, 2.0, 3.0
    iu, iv = np.meshgrid(np.arange(tw), np.arange(th))
    x = x0 + (iu.ravel() + 0.5) / tw * (x1 - x0)
    z = z0 + (iv.ravel() + 0.5) / th * (z1 - z0)
    pts = np.stack([x, np.full_like(x, 0.3), z], 1)  # y plays no part
    got = _texel_lookup(0, [x0, z0, x1, z1], pts, tw, th, fp64)
    np.testing.assert_array_equal(got, iv.ravel() * tw + iu.ravel())
    # outside the rectangle: clamped to the edge texels
    outside = np.array([[x0 - 1.0, 0.0, z0 - 1.0], [x1 + 1.0, 0.0, z1 + 1.0], [x1 + 1.0, 0.0, z0 - 1.0]])
    np.testing.assert_array_equal(_texel_lookup(0, [x0, z0, x1, z1], outside, tw, th, fp64),
                                  [0, th * tw - 1, tw - 1])


@pytest.mark.parametrize("fp64", [False, True])
def test_spherical_texel_lookup_hits_texel_centres(fp64):
    # column iu <-> azimuth phi = atan2(q.x, q.z) in [-pi, pi), row iv <-> polar
    # angle theta = acos(q.y) in [0, pi]: a point at radius 2.5 about the centre
    # in direction (sin t sin p, cos t, sin t cos p) at a texel centre's angles
    tw, th = 8, 6
    c = np.array([0.3, -0.7, 1.1])
    iu, iv = np.meshgrid(np.arange(tw), np.arange(th))
    phi = -np.pi + (iu.ravel() + 0.5) / tw * 2.0 * np.pi
    theta = (iv.ravel() + 0.5) / th * np.pi
    q = np.stack([np.sin(theta) * np.sin(phi), np.cos(theta), np.sin(theta) * np.cos(phi)], 1)
    got = _texel_lookup(1, [c[0], c[1], c[2], 0.0], c + 2.5 * q, tw, th, fp64)
    np.testing.assert_array_equal(got, iv.ravel() * tw + iu.ravel())
    # the poles: top row and bottom row
    poles = np.array([c + [0.0, 1.0, 0.0], c + [0.0, -1.0, 0.0]])
    got = _texel_lookup(1, [c[0], c[1], c[2], 0.0], poles, tw, th, fp64)
    assert got[0] // tw == 0 and got[1] // tw == th - 1
