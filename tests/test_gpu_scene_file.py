"""GPU test of the JSON scene loader (orf_scene_load_json; restates
ORF/OptixSceneLoader.cpp:105-177 and the readJSON / loadFromJSON of the classes
it dispatches to) through the C-ABI.

A scene file in the reference's schema (scenes/jsonscene.json: Integrator,
Background, Camera, Lights, Geometries with mtl_name / mtl_parameters) with OBJ
meshes and a Radiance HDR environment written here is loaded, rendered, and
compared per sample with

- the oracle rendering the same scene written out by hand in world space, and
- a context built through the C-ABI calls from that hand-written scene.

The hand-written world-space geometry uses quarter and half turns, whose images
are signed permutations of the coordinates worked out below, so the test pins
the object transform T * Rz * Ry * Rx * R(axis, angle) * S
(ORF/Geometry.cpp:286-291) and the inverse-transpose normals instead of
restating them.  Tolerance: |gpu - oracle| <= 1e-5 |oracle| + 1e-6 on >= 99% of
samples (shared FP32 transcendentals; a branch flipped on a last bit diverges);
loaded vs hand-built context (quarter turns in FP64 leave 1e-17 residues that
can flip a branch): the same bar on >= 99.5% of samples.
"""
import json
import sys

import numpy as np
import pytest

from open_rendering_framework_amd import rt, scenes
from tests import rgbe

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-5, 1e-6
W, H = 48, 40


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _write_obj(path, verts, tris, normals=None):
    lines = ["# written by the scene-file test"]
    lines += ["v %r %r %r" % tuple(float(c) for c in v) for v in verts]
    if normals is not None:
        lines += ["vn %r %r %r" % tuple(float(c) for c in n) for n in normals]
        lines += ["f %d//%d %d//%d %d//%d" % (a + 1, a + 1, b + 1, b + 1, c + 1, c + 1) for a, b, c in tris]
    else:
        lines += ["f %d %d %d" % (a + 1, b + 1, c + 1) for a, b, c in tris]
    path.write_text("\n".join(lines) + "\n")


def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32)


def _scene_files(tmp_path, translucent=False):
    """Writes the scene (JSON, OBJs, HDR) and returns (json path, equivalent SceneDesc)."""
    # ground: unit quad written as one polygon (a fan), scaled to 20 x 20 at y = -1
    gv = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    (tmp_path / "plane.obj").write_text(
        "v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nvn 0 1 0\nf 1//1 4//1 3//1 2//1\n")
    ground = scenes.Mesh(_f32(gv * [10, 1, 10] + [0, -1, 0]), np.array([[0, 3, 2], [0, 2, 1]], np.int32), 0,
                         normals=np.array([[0, 1, 0]], np.float32), nindex=np.zeros((2, 3), np.int32))
    # ball: scale (0.5, 1, 1), y_rotation 90, position (0.2, 0, 0)
    #   Ry(90) (x, y, z) = (z, y, -x):  p -> (z, y, -0.5 x) + (0.2, 0, 0)
    #   normals by Ry S^-1:             n -> (nz, ny, -2 nx)
    bv, bn, bt = scenes.icosphere(3)
    _write_obj(tmp_path / "ball.obj", bv, bt, bn)
    x, y, z = bv.T.astype(np.float64)
    ball = scenes.Mesh(_f32(np.stack([z + 0.2, y, -0.5 * x], 1)), bt, 1,
                       normals=_f32(np.stack([bn[:, 2], bn[:, 1], -2.0 * bn[:, 0].astype(np.float64)], 1)))
    # gem: uniform scale 0.6, rotation [180, 0, 1, 0], x_rotation 90, z_rotation 90, position (-1.2, -0.4, 0.3)
    #   R(180, y) (x, y, z) = (-x, y, -z);  Rx(90) (x, y, z) = (x, -z, y);  Rz(90) (x, y, z) = (-y, x, z)
    #   Rz Rx R (x, y, z) = Rz Rx (-x, y, -z) = Rz (-x, z, y) = (-z, -x, y)
    #   normals by (Rz Rx R) S^-1:  n -> (-nz, -nx, ny) / 0.6
    gv2, gn2, gt2 = scenes.icosphere(2, seed=3, amplitude=0.1)
    _write_obj(tmp_path / "gem.obj", gv2, gt2, gn2)
    x, y, z = gv2.T.astype(np.float64)
    gem = scenes.Mesh(_f32(0.6 * np.stack([-z, -x, y], 1) + [-1.2, -0.4, 0.3]), gt2, 2,
                      normals=_f32(np.stack([-gn2[:, 2], -gn2[:, 0], gn2[:, 1]], 1).astype(np.float64) / 0.6))
    # light: 1 x 1 quad facing -y, moved to y = 3
    lv, lt = scenes.quad(0.0, 0.5, facing_down=True)
    ln = np.tile([[0, -1, 0]], (4, 1)).astype(np.float32)
    _write_obj(tmp_path / "light.obj", lv, lt, ln)
    lw = _f32(lv.astype(np.float64) + [0, 3, 0])
    light_mesh = scenes.Mesh(lw, lt, 3, normals=ln)
    # sky: RGBE image; the environment is its decoded values, bottom row first
    sky = scenes.sky_envmap(64, 32, peak=2e3)[..., :3]
    enc = rgbe.to_rgbe(sky[::-1])
    rgbe.write(tmp_path / "sky.hdr", enc)
    env = np.ones((32, 64, 4), np.float32)
    env[..., :3] = rgbe.decode(enc)[::-1]

    if translucent:
        gem_mtl = {"mtl_name": "translucent_shader",
                   "mtl_parameters": {"ior": 1.3, "absorption": [0.0011, 0.0024, 0.014],
                                      "scattering": [2.55, 3.21, 3.77], "meancosine": [0.0, 0.0, 0.0],
                                      "scale": 100.0, "dipole_model": 1}}
        gem_desc = scenes.Material("subsurface_scattering_shader.cu",
                                   floats={"ior": (1.3,), "absorption": (0.0011, 0.0024, 0.014),
                                           "scattering": (2.55, 3.21, 3.77), "meancosine": (0.0, 0.0, 0.0),
                                           "scale": (100.0,)}, uints={"dipole_model": (1,)})
    else:
        gem_mtl = {"mtl_name": "rough_transparent_shader",
                   "mtl_parameters": {"ior": 1.5, "roughness": 0.2, "microfacet_model": 1, "normal_distribution": 1,
                                      "glass_absorption": [0.2, 0.1, 0.05]}}
        gem_desc = scenes.Material("rough_transparent_shader.cu",
                                   floats={"ior": (1.5,), "roughness": (0.2,), "glass_absorption": (0.2, 0.1, 0.05)},
                                   uints={"microfacet_model": (1,), "normal_distribution": (1,)})
    doc = {
        "frame_count": 16,
        "buffer_path": "out/film.raw",
        "Integrator": {"type": "PathTracer", "parameters": {"max_depth": 5, "scene_epsilon": 1e-4,
                                                            "exception_color": [1.0, 0.0, 1.0]}},
        "Background": {"type": "EnvMapBackground", "parameters": {"path": "sky.hdr",
                                                                  "background_color": [0.0, 0.0, 0.0]}},
        "Camera": {"type": "PinholeCamera", "parameters": {"eye": [0.0, 1.5, 4.0], "lookat": [0.0, 0.0, 0.0],
                                                           "up": [0.0, 1.0, 0.0], "vfov": 40, "width": W,
                                                           "height": H}},
        "Lights": [
            {"type": "TrianglesAreaLight", "parameters": {"path": "light.obj", "position": [0.0, 3.0, 0.0],
                                                          "radiance": [10.0, 10.0, 10.0]}},
            {"type": "PointLight", "parameters": {"radiance": [5.0, 4.0, 3.0], "position": [2.0, 4.0, 2.0]}},
            {"type": "DirectionalLight", "parameters": {"radiance": [0.5, 0.5, 0.6], "direction": [-1.0, -1.0, -1.0]}},
        ],
        "Geometries": [
            {"type": "obj", "parameters": {"path": "plane.obj", "position": [0.0, -1.0, 0.0], "scale": [10.0, 1.0, 10.0],
                                           "material": {"mtl_name": "diffuse_shader",
                                                        "mtl_parameters": {"diffuse_color": [0.5, 0.5, 0.5]}}}},
            {"type": "obj", "parameters": {"path": "ball.obj", "position": [0.2, 0.0, 0.0], "scale": [0.5, 1.0, 1.0],
                                           "y_rotation": 90,
                                           "material": {"mtl_name": "metallic_shader",
                                                        "mtl_parameters": {"real_ior": [0.27, 0.27, 0.27],
                                                                           "imaginary_ior": [2.78, 2.78, 2.78],
                                                                           "roughness": [0.3],
                                                                           "microfacet_model": 1,
                                                                           "normal_distribution": 1}}}},
            {"type": "obj", "parameters": {"path": "gem.obj", "position": [-1.2, -0.4, 0.3], "scale": [0.6, 0.6, 0.6],
                                           "rotation": [180, 0, 1, 0], "x_rotation": 90, "z_rotation": 90,
                                           "material": gem_mtl}},
        ],
    }
    (tmp_path / "scene.json").write_text(json.dumps(doc, indent=2))

    mats = [scenes.Material("diffuse_shader.cu", floats={"diffuse_color": (0.5, 0.5, 0.5)}),
            scenes.Material("metallic_shader.cu", floats={"roughness": (0.3, 0.3)},
                            uints={"microfacet_model": (1,), "normal_distribution": (1,)},
                            ior_complex=((0.27, 2.78),) * 3),
            gem_desc,
            scenes.Material("arealight_shader.cu", floats={"radiance": (10.0, 10.0, 10.0)})]
    lights = [scenes.Light("triangles", (10.0, 10.0, 10.0), tri_verts=lw[lt].reshape(-1, 9),
                           tri_normals=ln[lt].reshape(-1, 9)),
              scenes.Light("point", (5.0, 4.0, 3.0), position=(2.0, 4.0, 2.0)),
              scenes.Light("directional", (0.5, 0.5, 0.6), direction=(-1.0, -1.0, -1.0))]
    eye, U, V, Wv = scenes.pinhole((0.0, 1.5, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, W, H)
    desc = scenes.SceneDesc([ground, ball, gem, light_mesh], mats, lights, env, True, eye, U, V, Wv, W, H, 5, 1e-4)
    return tmp_path / "scene.json", desc


def _ok(got, want, rtol=RTOL, atol=ATOL):
    return (np.abs(got - want) <= rtol * np.abs(want) + atol).all(axis=-1)


def test_loaded_scene_matches_hand_written_scene(tmp_path):
    from tests.oracle_scene import OracleScene

    path, desc = _scene_files(tmp_path)
    ctx = rt.Context()
    info = ctx.load_scene(path, env_importance=True)
    _say("info", info)
    assert (info["width"], info["height"], info["frame_count"]) == (W, H, 16)
    assert (info["meshes"], info["materials"], info["lights"]) == (4, 4, 3)
    assert info["buffer_path"] == str(tmp_path / "out/film.raw")
    ctx.accel_build()
    hand = scenes.to_context(desc)
    hand.accel_build()
    orc = OracleScene(desc)
    ys, xs = np.mgrid[0:H, 0:W]
    for frame in (0, 5):
        _say("frame", frame)
        got = []
        for c in (ctx, hand):
            c.film_clear()
            c.launch(W, H, frame, 1)
            got.append(c.film_sums(W, H)[..., :3])
        want = orc.samples(W, H, xs.ravel(), ys.ravel(), frame).reshape(H, W, 3)
        same = _ok(got[0], got[1])
        ok = _ok(got[0], want)
        _say("  loaded vs hand-built %.5f, loaded vs oracle %.5f, mean %.4g / %.4g"
             % (same.mean(), ok.mean(), got[0].mean(), want.mean()))
        assert same.mean() >= 0.995
        assert ok.mean() >= 0.99
        assert want.mean() > 0.05  # the scene is lit


def test_loaded_translucent_scene(tmp_path):
    path, desc = _scene_files(tmp_path, translucent=True)
    ctx = rt.Context()
    ctx.load_scene(path)
    hand = scenes.to_context(desc)
    a, b = ctx.scattering_properties(2), hand.scattering_properties(2)
    for k in a:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)
    ctx.accel_build()
    p = ctx.probes()
    _say("probes", p["count"], "rmin %.5f" % p["rmin"])
    assert 1 << 15 < p["count"] <= 1 << 16  # the loader's default probe set of 2^16


@pytest.mark.parametrize("patch,match", [
    (lambda d: d["Geometries"][0]["parameters"]["material"].update(mtl_name="anisotropic_structure_shader"),
     "anisotropic_structure_shader"),
    (lambda d: d["Lights"].append({"type": "SpotLight", "parameters": {}}), "SpotLight"),
    (lambda d: d["Camera"].update(type="EnvCamera"), "EnvCamera"),
    (lambda d: d["Geometries"][1]["parameters"].update(path="missing.obj"), "missing.obj"),
])
def test_unsupported_or_broken_scene_is_refused(tmp_path, patch, match):
    path, _ = _scene_files(tmp_path)
    doc = json.loads(path.read_text())
    patch(doc)
    path.write_text(json.dumps(doc))
    with pytest.raises(rt.OrfError, match=match):
        rt.Context().load_scene(path)


def test_malformed_json_is_refused(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"Camera": {"type": "PinholeCamera", }')
    with pytest.raises(rt.OrfError):
        rt.Context().load_scene(p)


def test_missing_hdr_falls_back_to_background_color(tmp_path):
    """loadHDRTexture's fallback (ORF/HDRLoader.cpp:243-259): a 1 x 1 environment of background_color."""
    path, _ = _scene_files(tmp_path)
    doc = json.loads(path.read_text())
    doc["Background"]["parameters"].update(path="nothing.hdr", background_color=[0.25, 0.5, 0.75])
    doc["Camera"]["parameters"].update(lookat=[0.0, 3.0, -4.0])  # top rows look past the light into the sky
    path.write_text(json.dumps(doc))
    ctx = rt.Context()
    ctx.load_scene(path)
    ctx.accel_build()
    ctx.film_clear()
    ctx.launch(W, H, 0, 1)
    top = ctx.film_sums(W, H)[H - 1, :, :3]
    np.testing.assert_allclose(top, np.tile([0.25, 0.5, 0.75], (W, 1)), rtol=1e-6, atol=0)
