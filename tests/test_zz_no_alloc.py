"""orf_launch_noalloc (allocation-free launch into a caller-owned film) against orf_launch.

Two calls in a row, the second on a smaller film, each compared with orf_launch
on the same scene, frames and spp: sample counts and ray counters exactly, the
FP64 RGB sums to 1e-12 relative (both add FP64 partial sums with atomics whose
order is not fixed).  Then the refusals: a scene change not yet on the device,
and the depth integrator.
"""
import sys

import numpy as np
import pytest
import torch

from open_rendering_framework_amd import rt, scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _reference(ctx, w, h, frame0, spp):
    ctx.film_clear()
    ctx.launch(w, h, frame0, spp)
    return ctx.film_sums(w, h), ctx.counters()


def _noalloc(ctx, w, h, frame0, spp, with_counters=True):
    film = torch.zeros(h * w * 4, dtype=torch.float64, device="cuda")
    cnt = torch.zeros(4, dtype=torch.int64, device="cuda") if with_counters else None
    ctx.launch_noalloc(w, h, frame0, spp, film, cnt)
    torch.cuda.synchronize()
    c = None
    if cnt is not None:
        c = dict(zip(("nonfinite", "ext_rays", "shadow_rays", "paths"), [int(x) for x in cnt.cpu().tolist()]))
    return film.cpu().numpy().reshape(h, w, 4), c


def _compare(tag, got, want, cg, cw):
    assert np.array_equal(got[..., 3], want[..., 3]), tag
    np.testing.assert_allclose(got[..., :3], want[..., :3], rtol=1e-12, atol=1e-12)
    if cg is not None:
        assert cg == cw, (tag, cg, cw)
    _say("%s: counts equal, max |d| %.3g, counters %s" % (tag, np.abs(got - want).max(), cg))


def test_noalloc_matches_launch_twice_then_smaller():
    d = scenes.path_scene(width=64, height=64, level=4, env_size=(256, 128))
    ctx = scenes.to_context(d)
    ctx.accel_build()
    want, cw = _reference(ctx, 64, 64, 3, 4)
    got, cg = _noalloc(ctx, 64, 64, 3, 4)
    _compare("64x64 frame 3 spp 4", got, want, cg, cw)
    want, cw = _reference(ctx, 32, 32, 7, 2)
    got, cg = _noalloc(ctx, 32, 32, 7, 2)
    _compare("32x32 frame 7 spp 2", got, want, cg, cw)
    got, _ = _noalloc(ctx, 32, 32, 7, 2, with_counters=False)
    _compare("32x32 without counters", got, want, None, None)


def test_noalloc_refusals():
    d = scenes.path_scene(width=16, height=16, level=2, env_size=(64, 32))
    ctx = scenes.to_context(d)
    film = torch.zeros(16 * 16 * 4, dtype=torch.float64, device="cuda")
    with pytest.raises(rt.OrfError, match="orf_accel_build"):
        ctx.launch_noalloc(16, 16, 0, 1, film)  # scene never built
    ctx.accel_build()
    ctx.launch_noalloc(16, 16, 0, 1, film)
    torch.cuda.synchronize()
    assert float(film.view(16, 16, 4)[..., 3].sum()) == 256.0
    ctx.integrator(rt.INTEGRATOR_DEPTH)
    with pytest.raises(rt.OrfError, match="path tracer"):
        ctx.launch_noalloc(16, 16, 0, 1, film)
