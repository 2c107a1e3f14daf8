"""GPU parity of the disk and sphere lights (NEE: ORF/LightSampler.h:73-141;
emitters seen by rays: CU/disk_mesh.cu:41-63, CU/sphere_mesh.cu:42-91 with the
arealight material) through the C-ABI against the CPU oracle.

Tolerances: per sample |gpu - oracle| <= 1e-5 |oracle| + 1e-6 on >= 99% of
samples (shared FP32 transcendentals; a branch flipped on a last bit diverges);
converged image vs the FP64 oracle: relative RMSE < 3e-2; the wavefront and
persistent batch forms vs the megakernel: relative 1e-9 (order of FP64 atomic
additions), equal ray counts.
"""
import sys

import numpy as np
import pytest

from open_rendering_framework_amd import rt, scenes

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-5, 1e-6
W = 48


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _scene():
    """The path-tracer scene at test scale with a disk light (tilted towards the
    camera and down) and a sphere light added next to its triangle light and
    environment, both in view."""
    d = scenes.path_scene(width=W, height=W, level=4, env_size=(128, 64), max_depth=6)
    d.lights.append(scenes.Light("disk", (6.0, 5.0, 4.0), position=(1.2, 1.4, 0.6), radius=0.35, theta=110.0,
                                 phi=340.0))
    d.lights.append(scenes.Light("sphere", (3.0, 4.0, 6.0), position=(-1.4, 0.2, 0.9), radius=0.3))
    return d


def _ok(got, want):
    return (np.abs(got - want) <= RTOL * np.abs(want) + ATOL).all(axis=-1)


def test_per_sample_parity():
    from tests.oracle_scene import OracleScene

    d = _scene()
    ctx = scenes.to_context(d)
    ctx.accel_build()
    orc = OracleScene(d)
    ys, xs = np.mgrid[0:W, 0:W]
    for frame in (0, 7):
        ctx.film_clear()
        ctx.launch(W, W, frame, 1)
        got = ctx.film_sums(W, W)[..., :3]
        want = orc.samples(W, W, xs.ravel(), ys.ravel(), frame).reshape(W, W, 3)
        ok = _ok(got, want)
        _say("frame %d samples ok %.5f mean %.4g / %.4g" % (frame, ok.mean(), got.mean(), want.mean()))
        assert ok.mean() >= 0.99


def test_emitters_in_view():
    """Camera rays that hit the emitters directly return their radiance: pixel
    sums equal to a light's radiance exist for both lights."""
    d = _scene()
    d.max_depth = 0
    ctx = scenes.to_context(d)
    ctx.accel_build()
    ctx.film_clear()
    ctx.launch(W, W, 0, 1)
    img = ctx.film_sums(W, W)[..., :3].reshape(-1, 3)
    for rad in ((6.0, 5.0, 4.0), (3.0, 4.0, 6.0)):
        n = int(np.all(img == np.float32(rad), axis=1).sum())
        _say("pixels showing", rad, n)
        assert n >= 4


def test_converged_vs_fp64_oracle():
    from tests.oracle_scene import OracleScene

    d = _scene()
    ctx = scenes.to_context(d)
    ctx.accel_build()
    spp = 32
    ctx.film_clear()
    ctx.launch(W, W, 0, spp)
    img = ctx.film(W, W)[..., :3]
    c = ctx.counters()
    ref, _ = OracleScene(d, fp64=True).render(W, W, 0, spp)
    ref = ref[..., :3] / ref[..., 3:4]
    rel = np.sqrt(((img - ref) ** 2).mean()) / ref.mean()
    _say("rel rmse vs fp64 %.3g, counters %s" % (rel, c))
    assert c["nonfinite"] == 0
    assert rel < 3e-2


def test_batch_forms_match_megakernel():
    d = _scene()
    ctx = scenes.to_context(d)
    items = [{"width": W, "height": W, "spp": 4, "frame0": f, "eye": d.eye, "U": d.U, "V": d.V, "W": d.W}
             for f in (0, 11)]
    ctx.render_batch(items, None, mode=rt.BATCH_MEGAKERNEL)
    ref = [ctx.batch_film(i, W, W, sums=True) for i in range(len(items))]
    ref_counts = ctx.batch_counters()
    for mode in (rt.BATCH_WAVEFRONT, rt.BATCH_PERSISTENT):
        ctx.render_batch(items, None, mode=mode)
        counts = ctx.batch_counters()
        _say("mode %d counters %s (megakernel %s)" % (mode, counts, ref_counts))
        for i in range(len(items)):
            a = ctx.batch_film(i, W, W, sums=True)
            np.testing.assert_allclose(a[..., :3], ref[i][..., :3], rtol=1e-9, atol=1e-12)
        assert counts["ext_rays"] == ref_counts["ext_rays"]
        assert counts["shadow_rays"] == ref_counts["shadow_rays"]


def test_bad_radius_is_refused():
    ctx = rt.Context()
    with pytest.raises(rt.OrfError, match="radius"):
        ctx.disk_light((0, 0, 0), 0.0, 0.0, 0.0, (1, 1, 1))
    with pytest.raises(rt.OrfError, match="radius"):
        ctx.sphere_light((0, 0, 0), float("inf"), (1, 1, 1))


def test_differentiable_form_refuses_emitters():
    import torch

    d = _scene()
    ctx = scenes.to_context(d)
    desc = rt.DiffDesc.make(scenes.DIFF_ALBEDO_MATERIAL, scenes.DIFF_ROUGHNESS_MATERIAL, (4, 4),
                            scenes.DIFF_ALBEDO_UV, scenes.DIFF_ROUGHNESS_UV)
    views = scenes.ring_views(n=1, size=16, spp=1)
    a = torch.full((4, 4, 3), 0.5, device="cuda")
    r = torch.full((4, 4), 0.3, device="cuda")
    with pytest.raises(rt.OrfError, match="disk or sphere"):
        ctx.diff_render(desc, views, a, r)


def test_scene_file_disk_and_sphere_lights(tmp_path):
    """DiskAreaLight / SphericalLight entries of a scene file (ORF/Light.cpp:389-423,
    525-548) render per sample as the oracle on the hand-written equivalent."""
    import json

    from tests.oracle_scene import OracleScene
    from tests.test_gpu_scene_file import H as SH, W as SW, _scene_files

    path, desc = _scene_files(tmp_path)
    doc = json.loads(path.read_text())
    doc["Lights"] += [
        {"type": "DiskAreaLight", "parameters": {"position": [0.8, 1.2, 1.0], "radius": 0.3, "theta": 120.0,
                                                 "phi": 20.0, "radiance": [4.0, 3.0, 2.0]}},
        {"type": "SphericalLight", "parameters": {"position": [-0.9, 0.9, 0.6], "radius": 0.2,
                                                  "radiance": [2.0, 3.0, 5.0]}},
    ]
    path.write_text(json.dumps(doc))
    desc.lights += [scenes.Light("disk", (4.0, 3.0, 2.0), position=(0.8, 1.2, 1.0), radius=0.3, theta=120.0, phi=20.0),
                    scenes.Light("sphere", (2.0, 3.0, 5.0), position=(-0.9, 0.9, 0.6), radius=0.2)]
    ctx = rt.Context()
    info = ctx.load_scene(path, env_importance=True)
    assert info["lights"] == 5
    ctx.accel_build()
    ctx.film_clear()
    ctx.launch(SW, SH, 1, 1)
    got = ctx.film_sums(SW, SH)[..., :3]
    ys, xs = np.mgrid[0:SH, 0:SW]
    want = OracleScene(desc).samples(SW, SH, xs.ravel(), ys.ravel(), 1).reshape(SH, SW, 3)
    ok = _ok(got, want)
    _say("scene-file disk/sphere lights: samples ok %.5f" % ok.mean())
    assert ok.mean() >= 0.99
