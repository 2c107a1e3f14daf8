"""GPU checks of the progressive frame driver and the film files (OptixScene's
frame loop, saveScene and loadBuffer: ORF/OptixScene.cpp:139-168, 202-299;
OptixSceneLoader::saveJSONScene, ORF/OptixSceneLoader.cpp:48-70) through the C-ABI.

- per-frame driver: frames [0, 3) of the translucent scene, each after its own
  probe pass, equal the oracle's probe pass + samples per frame
  (|gpu - oracle| <= 1e-5 |oracle| + 1e-6 per pixel on >= 99% of pixels);
- .raw: width x height FP32 RGBA of the film's mean, launch-row order, alpha 1,
  byte-exact against the film read back;
- resume: frames [0, 4) saved, loaded as 4 frames, then frames [4, 8) ==
  frames [0, 8) in one context (relative 1e-6: the saved mean is FP32);
- PNG: decoded with zlib it equals clamp((int)(mean * 255)), top row first, and
  every chunk CRC checks;
- scene save: the loop with quit_and_save rewrites the document with
  frame_count / buffer_path and writes .png/.raw; reopening it loads the film
  (no frames rendered) and reproduces the image exactly.
"""
import json
import struct
import sys
import zlib

import numpy as np
import pytest

from open_rendering_framework_amd import rt, scenes

pytestmark = pytest.mark.gpu

W, H = 40, 32
MILK = 1


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _path_ctx():
    d = scenes.path_scene(width=W, height=H, level=3, env_size=(64, 32), max_depth=5)
    d.eye, d.U, d.V, d.W = scenes.pinhole((0.0, 1.5, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, W, H)
    return d, scenes.to_context(d)


def test_frames_with_per_frame_probe_pass_match_oracle():
    from tests.oracle_scene import OracleScene

    d = scenes.subsurface_scene(width=W, height=H, level=5, glass_level=3, env_size=(128, 64), probes=8192,
                                probe_samples=4)
    ctx = scenes.to_context(d)
    ctx.accel_build()
    orc = OracleScene(d)
    orc.probes_set(ctx.probes(), 4, MILK)
    ctx.film_clear()
    ctx.render_frames(W, H, 0, 3)
    got = ctx.film_sums(W, H)[..., :3]
    ys, xs = np.mgrid[0:H, 0:W]
    want = np.zeros((H, W, 3))
    for f in range(3):
        orc.probe_pass(f)
        want += orc.samples(W, H, xs.ravel(), ys.ravel(), f).reshape(H, W, 3)
    ok = (np.abs(got - want) <= 1e-5 * np.abs(want) + 1e-6).all(axis=-1)
    _say("per-frame driver vs oracle: %.5f of pixels, mean %.4g / %.4g" % (ok.mean(), got.mean(), want.mean()))
    assert ok.mean() >= 0.99


def test_raw_file_is_the_fp32_mean(tmp_path):
    _, ctx = _path_ctx()
    ctx.film_clear()
    ctx.render_frames(W, H, 0, 3)
    p = tmp_path / "film.raw"
    ctx.save_raw(p)
    assert p.stat().st_size == W * H * 16
    raw = np.fromfile(p, np.float32).reshape(H, W, 4)
    np.testing.assert_array_equal(raw, ctx.film(W, H).astype(np.float32))
    assert np.all(raw[..., 3] == 1.0)


def test_resume_from_raw_continues_the_running_mean(tmp_path):
    _, a = _path_ctx()
    a.film_clear()
    a.render_frames(W, H, 0, 4)
    p = tmp_path / "half.raw"
    a.save_raw(p)
    _, b = _path_ctx()
    b.load_raw(p, W, H, 4)
    np.testing.assert_array_equal(b.film(W, H).astype(np.float32), np.fromfile(p, np.float32).reshape(H, W, 4))
    b.render_frames(W, H, 4, 8)
    _, c = _path_ctx()
    c.film_clear()
    c.render_frames(W, H, 0, 8)
    fb, fc = b.film(W, H)[..., :3], c.film(W, H)[..., :3]
    rel = np.abs(fb - fc).max() / np.abs(fc).max()
    _say("resume vs straight 8 frames: max rel %.3g" % rel)
    np.testing.assert_allclose(fb, fc, rtol=1e-6, atol=1e-7 * np.abs(fc).max())
    with pytest.raises(rt.OrfError, match="FP32 RGBA"):
        b.load_raw(p, W + 1, H, 4)


def _png_decode(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    i, chunks = 8, {}
    while i < len(data):
        n, = struct.unpack(">I", data[i:i + 4])
        typ = data[i + 4:i + 8]
        body = data[i + 8:i + 8 + n]
        crc, = struct.unpack(">I", data[i + 8 + n:i + 12 + n])
        assert zlib.crc32(typ + body) & 0xFFFFFFFF == crc, typ
        chunks[typ] = chunks.get(typ, b"") + body
        i += 12 + n
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    assert (depth, ctype) == (8, 6)
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 4)


def test_png_matches_buffer_to_image(tmp_path):
    _, ctx = _path_ctx()
    ctx.film_clear()
    ctx.render_frames(W, H, 0, 2)
    p = tmp_path / "shot.png"
    ctx.save_png(p)
    img = _png_decode(p.read_bytes())
    mean = ctx.film(W, H).astype(np.float32)
    want = np.clip(np.trunc(mean * np.float32(255.0)), 0, 255).astype(np.uint8)[::-1]
    assert img.shape == (H, W, 4)
    np.testing.assert_array_equal(img, want)


def test_scene_save_and_reload(tmp_path):
    from open_rendering_framework_amd.progressive import ProgressiveScene
    from tests.test_gpu_scene_file import _scene_files

    path, _ = _scene_files(tmp_path)
    doc = json.loads(path.read_text())
    doc["frame_count"] = 4
    del doc["buffer_path"]
    path.write_text(json.dumps(doc))
    p = ProgressiveScene(path, quit_and_save=True, env_importance=True)
    n = 0
    while p.render_scene():
        n += 1
    _say("frames rendered", p.frame)
    assert p.frame == 4
    saved = json.loads(path.read_text())
    assert saved["frame_count"] == 4 and saved["buffer_path"] == "scene.raw"
    assert saved["Camera"] == doc["Camera"] and saved["Geometries"] == doc["Geometries"]
    assert (tmp_path / "scene.png").exists() and (tmp_path / "scene.raw").stat().st_size == 48 * 40 * 16
    first = p.image()
    q = ProgressiveScene(path, env_importance=True)
    assert q.max_frame == 4 and q.frame == 4
    assert not q.render_scene() and q.frame == 4  # nothing left to render
    np.testing.assert_array_equal(q.image(), first)
