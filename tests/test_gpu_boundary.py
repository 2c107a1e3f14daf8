"""GPU tests of the rest of the OptiX-host-API surface on the C-ABI: a compiled
C++ caller, user-format structs, Matrix4x4 variables, instance scope, per-ray-type
program binding and texture samplers.

- tests/cabi/orf_caller (C++, linked against liborf_rt.so, built by build())
  drives a scene the way the reference's classes drive OptiX: programs bound
  per ray type, ScatteringMaterialProperties as user data (computed in the
  caller as RoughTranslucentMaterial::computeCoefficients does), an instance
  Matrix4x4 and a transposed one, instance-scope emitter radiance, the raw
  light_buffer / triangle_light_buffer, an "envmap" texture sampler and an
  exception from a refused call.  Its film equals, pixel by pixel, the film
  of the same scene built through the plain calls (material variables,
  orf_mesh_set_transform, orf_light_add_*, orf_envmap_set): the sample counts
  exactly, the sums within 1e-5 relative on >= 99.5% of pixels (the caller's
  FP32 coefficients may differ from the library's in the last bit, which can
  move a rare Russian-roulette decision).
- User data: ScatteringMaterialProperties computed in numpy from the
  reference's formulas render per sample as the material variables do
  (>= 99.9% within 1e-5); a doubled global_coeff doubles the diffusion term.
- Instance scope: one mesh's instance-scope diffuse_color changes that mesh
  only (per-sample equal to a scene with two materials).
- Refusals: a triangle light entry past the triangle buffer, a wrong element
  size, an inverse_transform_matrix that is not the inverse, an unknown
  program, a CLAMP texture as envmap.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from open_rendering_framework_amd import rt, scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "cabi", "orf_caller")
W = H = 48
RTOL, ATOL = 1e-5, 1e-6


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _parts(level=3):
    gv, gt = scenes.quad(-1.0, 10.0, facing_down=False)
    sv, sn, st = scenes.icosphere(level)
    lv, lt = scenes.quad(3.0, 0.5, facing_down=True)
    env = scenes.sky_envmap(64, 32)
    cam = scenes.pinhole((0.0, 1.5, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, W, H)
    return gv, gt, sv, sn, st, lv, lt, env, cam


def _write_scene(path, spp):
    gv, gt, sv, sn, st, lv, lt, env, cam = _parts()
    hdr = np.array([W, H, spp, len(gv), len(gt), len(sv), len(st), env.shape[1], env.shape[0]], np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([c for v in cam for c in v], np.float32).tobytes())
        for a, t in ((gv, np.float32), (gt, np.int32), (sv, np.float32), (sn, np.float32), (st, np.int32),
                     (lv, np.float32), (lt, np.int32), (env, np.float32)):
            f.write(np.ascontiguousarray(a, t).tobytes())


def _milk_vars(ctx, h):
    ctx.set_float("absorption", 0.0011, 0.0024, 0.014, material=h)
    ctx.set_float("scattering", 2.55, 3.21, 3.77, material=h)
    ctx.set_float("meancosine", 0.0, 0.0, 0.0, material=h)
    ctx.set_float("scale", 100.0, material=h)
    ctx.set_float("ior", 1.3, material=h)
    ctx.set_uint("dipole_model", 0, material=h)


def _plain_context():
    """The caller's scene through the plain calls."""
    gv, gt, sv, sn, st, lv, lt, env, cam = _parts()
    ctx = rt.Context()
    ctx.set_int("max_depth", 8)
    ctx.set_float("scene_epsilon", 1e-4)
    for k, v in zip(("eye", "U", "V", "W"), cam):
        ctx.set_float(k, *v)
    ground = ctx.material("diffuse_shader.cu")
    ctx.set_float("diffuse_color", 0.5, 0.5, 0.5, material=ground)
    metal = ctx.material("metallic_shader.cu")
    ctx.set_float("roughness", 0.3, 0.3, material=metal)
    ctx.set_uint("microfacet_model", 1, material=metal)
    ctx.set_uint("normal_distribution", 1, material=metal)
    ctx.set_conductor_ior(metal, ((0.27, 2.78),) * 3)
    milk = ctx.material("subsurface_scattering_shader.cu")
    _milk_vars(ctx, milk)
    ctx.set_uint("probe_count", 4096)
    ctx.set_uint("probe_samples", 4)
    emitter = ctx.material("arealight_shader.cu")
    ctx.set_float("radiance", 10.0, 10.0, 10.0, material=emitter)
    ctx.mesh(gv, gt, ground)
    m = ctx.mesh(sv, st, metal, sn)
    s = ctx.mesh(sv, st, milk, sn)
    ctx.mesh(lv, lt, emitter)
    ctx.set_transform(m, [[1, 0, 0, -1.2], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])
    ctx.set_transform(s, [[1, 0, 0, 1.2], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])
    tv = lv[lt].reshape(-1, 9)
    tn = np.tile(np.array([0, -1, 0], np.float32), (len(lt), 3))
    ctx.triangle_light(tv, tn, (10.0, 10.0, 10.0))
    ctx.point_light((-2.0, 2.5, 1.5), (3.0, 2.5, 2.0))
    ctx.envmap(env, True)
    return ctx


def _frac_ok(got, want):
    ok = (np.abs(got - want) <= RTOL * np.abs(want) + ATOL).all(axis=-1)
    return ok.mean()


def test_compiled_caller_matches_plain_calls(tmp_path):
    assert os.path.exists(CALLER), "tests/cabi/orf_caller is built by build()"
    spp = 4
    scene, film = tmp_path / "scene.bin", tmp_path / "film.bin"
    _write_scene(scene, spp)
    _say("running", CALLER)
    r = subprocess.run([CALLER, str(scene), str(film)], capture_output=True, text=True, timeout=240)
    _say(r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0
    got = np.fromfile(film, np.float64).reshape(H, W, 4)
    ctx = _plain_context()
    ctx.film_clear()
    ctx.launch(W, H, 0, spp)
    want = ctx.film_sums(W, H)
    ctx.close()
    assert np.array_equal(got[..., 3], want[..., 3])
    frac = _frac_ok(got[..., :3] / spp, want[..., :3] / spp)
    _say("caller vs plain calls: %.5f of pixels within tolerance" % frac)
    assert frac >= 0.995
    assert np.isfinite(got).all() and got[..., :3].mean() > 0


def _reference_properties(sa, ss, g, scale, ior):
    """ScatteringMaterialProperties by ORF/RoughTranslucentMaterial.cpp:366-395 in FP32."""
    f = np.float32
    sa, ss, g = (np.asarray(x, f) * s for x, s in ((sa, f(scale)), (ss, f(scale)), (g, f(1))))

    def two_c1(n):
        n = f(n)
        if n >= 1:
            return f(-9.23372) + n * (f(22.2272) + n * (f(-20.9292) + n * (f(10.2291) + n * (f(-2.54396) + n * f(0.254913)))))
        return f(0.919317) + n * (f(-3.4793) + n * (f(6.75335) + n * (f(-7.80989) + n * (f(4.98554) - n * f(1.36881)))))

    def three_c2(n):
        n = f(n)
        if n >= 1:
            r = f(-1641.1) + n * (f(1213.67) + n * (f(-568.556) + n * (f(164.798) + n * (f(-27.0181) + n * f(1.91826)))))
            return r + (((f(135.926) / n) - f(656.175)) / n + f(1376.53)) / n
        return f(0.828421) + n * (f(-2.62051) + n * (f(3.36231) + n * (f(-1.95284) + n * (f(0.236494) + n * f(0.145787)))))

    p = np.zeros((), rt.SCATTERING_PROPERTIES)
    p["relative_ior"] = ior
    p["ior_real"] = ior
    p["absorption"], p["scattering"], p["meancosine"] = sa, ss, g
    red = ss * (f(1) - g)
    p["extinction"] = ss + sa
    p["reducedExtinction"] = red + sa
    p["deltaEddExtinction"] = ss * (f(1) - g * g) + sa
    D = f(1) / (f(3) * p["reducedExtinction"])
    p["D"] = D
    p["transport"] = np.sqrt(sa / D)
    cphi = f(0.25) * (f(1) - two_c1(ior))
    cphi_inv = f(0.25) * (f(1) - two_c1(f(1) / f(ior)))
    ce = f(0.5) * (f(1) - three_c2(ior))
    p["C_phi"], p["C_phi_inv"], p["C_E"] = cphi, cphi_inv, ce
    p["albedo"] = ss / p["extinction"]
    p["reducedAlbedo"] = red / p["reducedExtinction"]
    p["de"] = f(2.131) * D / np.sqrt(p["reducedAlbedo"])
    p["A"] = (f(1) - ce) / (f(2) * cphi)
    p["three_D"] = f(3) * D
    p["two_A_de"] = f(2) * p["A"] * p["de"]
    p["global_coeff"] = f(1) / (f(4) * cphi_inv) * f(1) / (f(4) * f(np.pi) * f(np.pi))
    p["one_over_three_ext"] = f(1) / (f(3) * p["extinction"])
    p["mean_transport"] = p["transport"].mean()
    return p


def _milk_scene(props=None):
    d = scenes.path_scene(width=W, height=H, level=3, env_size=(64, 32))
    d.materials.append(scenes.Material("subsurface_scattering_shader.cu"))
    d.meshes[1].material = len(d.materials) - 1
    d.ctx_uints = {"probe_count": (4096,), "probe_samples": (4,)}
    ctx = scenes.to_context(d)
    h = len(d.materials) - 1
    if props is None:
        _milk_vars(ctx, h)
    else:
        ctx.set_user_data("scattering_properties", props.tobytes(), material=h)
    return ctx


def test_user_data_scattering_properties():
    props = _reference_properties((0.0011, 0.0024, 0.014), (2.55, 3.21, 3.77), (0.0, 0.0, 0.0), 100.0, 1.3)
    a = _milk_scene()
    b = _milk_scene(props)
    lib = np.array(a.scattering_properties(3)["global_coeff"])
    np.testing.assert_allclose(props["global_coeff"], lib, rtol=1e-6)
    for ctx in (a, b):
        ctx.film_clear()
        ctx.launch(W, H, 2, 1)
    fa, fb = a.film_sums(W, H)[..., :3], b.film_sums(W, H)[..., :3]
    frac = _frac_ok(fb, fa)
    _say("user-data properties vs variables: %.5f within tolerance" % frac)
    assert frac >= 0.999
    props2 = props.copy()
    props2["global_coeff"] = props["global_coeff"] * 2
    c = _milk_scene(props2)
    for ctx in (a, c):
        ctx.film_clear()
        ctx.launch(W, H, 0, 16)
    ia, ic = a.film(W, H)[..., :3], c.film(W, H)[..., :3]
    assert ic.mean() > ia.mean() * 1.02  # the diffusion term doubles; the surface terms stay


def test_instance_scope_overrides_one_mesh():
    d = scenes.path_scene(width=W, height=H, level=3, env_size=(64, 32))
    d.meshes[1].material = 0  # the sphere shares the ground's Lambert material
    a = scenes.to_context(d)
    a.set_float("diffuse_color", 0.9, 0.2, 0.1, instance=1)
    d2 = scenes.path_scene(width=W, height=H, level=3, env_size=(64, 32))
    d2.materials.append(scenes.Material("diffuse_shader.cu", floats={"diffuse_color": (0.9, 0.2, 0.1)}))
    d2.meshes[1].material = len(d2.materials) - 1
    b = scenes.to_context(d2)
    for ctx in (a, b):
        ctx.film_clear()
        ctx.launch(W, H, 1, 1)
    frac = _frac_ok(a.film_sums(W, H)[..., :3], b.film_sums(W, H)[..., :3])
    _say("instance-scope colour vs two materials: %.5f" % frac)
    assert frac >= 0.999


def test_light_buffers_and_refusals():
    d = scenes.path_scene(width=16, height=16, level=2, env_size=(64, 32))
    ctx = scenes.to_context(d)
    L = np.zeros(1, rt.LIGHT_STRUCT)
    L["light_type"] = 2
    L["emitted_radiance"] = 10.0
    L["light_parameters"] = (2.0, 5.0, 6.0)  # triangles 5..6 of a 2-entry triangle buffer
    ctx.set_buffer("light_buffer", L)
    with pytest.raises(rt.OrfError, match="triangle_light_buffer"):
        ctx.launch(16, 16, 0, 1)
    L["light_parameters"] = (2.0, 0.0, 1.0)
    ctx.set_buffer("light_buffer", L)
    ctx.launch(16, 16, 0, 1)
    with pytest.raises(rt.OrfError, match="40 bytes"):
        ctx.set_buffer("light_buffer", np.zeros(2, np.float64))
    with pytest.raises(rt.OrfError, match="not the inverse"):
        ctx.set_matrix4x4("inverse_transform_matrix", np.eye(4) * 2, instance=1)
    with pytest.raises(rt.OrfError, match="no radiance closest hit"):
        ctx.material_program(0, rt.RAY_RADIANCE, "no_such_shader.cu", "closest_hit")
    tex = ctx.texture(np.ones((8, 16, 4), np.float32), wrap=rt.WRAP_CLAMP_TO_EDGE)
    with pytest.raises(rt.OrfError, match="REPEAT"):
        ctx.set_texture("envmap", tex)
