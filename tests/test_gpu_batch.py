"""GPU checks of batched renders (through the C-ABI).

The wavefront form runs shade_vertex() -- the megakernel's own shading code --
on queued hits, so for every item its film must equal the per-item megakernel
film up to the order of FP64 atomic additions (relative 1e-9), for every
wavefront configuration; ray counts must match exactly.  The persistent batched
megakernel runs the same path loop over different sample chunks: same bar.  Item images are also
compared with the FP64 oracle (relative RMSE stated below).
"""
import sys

import numpy as np
import pytest

from open_rendering_framework_amd import rt, scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a)
    sys.stdout.flush()


def _small_batch(n=12):
    d, items, remap = scenes.batch_workload(n=n, level=4, env_size=(256, 128), spp_scale=1 / 16, size_scale=0.25)
    return d, items, remap


def _films(ctx, items, sums=True):
    return [ctx.batch_film(i, it["width"], it["height"], sums=sums) for i, it in enumerate(items)]


@pytest.fixture(scope="module")
def batch():
    d, items, remap = _small_batch()
    ctx = scenes.to_context(d)
    ctx.render_batch(items, remap, mode=rt.BATCH_MEGAKERNEL)
    ref = _films(ctx, items)
    ref_counts = ctx.batch_counters()
    return d, items, remap, ctx, ref, ref_counts


CONFIGS = [
    dict(block_size=128, pool_size=1 << 16, sort_policy=1, regen_threshold=1.0, use_graph=1, graph_pairs=4),
    dict(block_size=64, pool_size=1 << 14, sort_policy=0, regen_threshold=0.5, use_graph=1, graph_pairs=1),
    dict(block_size=256, pool_size=1 << 15, sort_policy=1, regen_threshold=0.5, use_graph=0, graph_pairs=2),
]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_wavefront_equals_per_item_megakernel(batch, cfg):
    d, items, remap, ctx, ref, ref_counts = batch
    c = rt.WavefrontConfig()
    for k, v in cfg.items():
        setattr(c, k, v)
    secs = ctx.render_batch(items, remap, mode=rt.BATCH_WAVEFRONT, config=c)
    got = _films(ctx, items)
    counts = ctx.batch_counters()
    _say("cfg %s: %.4f s, counters %s (megakernel %s)" % (cfg, secs, counts, ref_counts))
    for i, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a[..., 3], b[..., 3]), i
        np.testing.assert_allclose(a[..., :3], b[..., :3], rtol=1e-9, atol=1e-12)
    assert counts["ext_rays"] == ref_counts["ext_rays"]
    assert counts["shadow_rays"] == ref_counts["shadow_rays"]
    assert counts["paths"] == ref_counts["paths"] == sum(it["width"] * it["height"] * it["spp"] for it in items)


def test_persistent_equals_per_item_megakernel(batch):
    d, items, remap, ctx, ref, ref_counts = batch
    secs = ctx.render_batch(items, remap, mode=rt.BATCH_PERSISTENT)
    got = _films(ctx, items)
    counts = ctx.batch_counters()
    _say("persistent: %.4f s, counters %s (megakernel %s)" % (secs, counts, ref_counts))
    for i, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a[..., 3], b[..., 3]), i
        np.testing.assert_allclose(a[..., :3], b[..., :3], rtol=1e-9, atol=1e-12)
    assert counts["ext_rays"] == ref_counts["ext_rays"]
    assert counts["shadow_rays"] == ref_counts["shadow_rays"]
    assert counts["paths"] == ref_counts["paths"]


def test_batch_items_vs_fp64_oracle(batch):
    from tests.oracle_scene import OracleScene

    d, items, remap, ctx, ref, _ = batch
    ctx.render_batch(items, remap, mode=rt.BATCH_WAVEFRONT)
    for i in range(3):
        it = items[i]
        dd = scenes.batch_workload(n=12, level=4, env_size=(256, 128), spp_scale=1 / 16, size_scale=0.25)[0]
        dd.meshes[1].material = int(remap[i][1])
        dd.eye, dd.U, dd.V, dd.W = it["eye"], it["U"], it["V"], it["W"]
        w, h = it["width"], it["height"]
        img = ctx.batch_film(i, w, h)[..., :3]
        orc, _ = OracleScene(dd, fp64=True).render(w, h, it["frame0"], it["spp"])
        orc = orc[..., :3] / orc[..., 3:4]
        rel = np.sqrt(((img - orc) ** 2).mean()) / orc.mean()
        _say("item %d (%s, %dx%d, %d spp): rel rmse vs fp64 oracle %.3g" % (i, it["material"], w, h, it["spp"], rel))
        assert rel < 2.5e-3  # measured at most 1.0e-3


def test_autotune_returns_valid_config(batch):
    d, items, remap, ctx, ref, _ = batch
    best, report = ctx.autotune(items, remap, spp_divisor=2)
    _say("autotune best %s" % best.as_dict())
    assert best.block_size in (64, 128, 256)
    assert len(report["trials"]) >= 5
    ctx.render_batch(items, remap, mode=rt.BATCH_WAVEFRONT, config=best)
    got = _films(ctx, items)
    for a, b in zip(got, ref):
        np.testing.assert_allclose(a[..., :3], b[..., :3], rtol=1e-9, atol=1e-12)


def test_full_batch_workload_forms_agree():
    """The benchmark's batch at full size (64 renders, 64x64 .. 1024x1024, 4 ..
    256 spp, level-7 sphere, 1024x512 env): wavefront and persistent films equal
    the per-item megakernel films on every item (relative 1e-9) with equal ray
    counts."""
    d, items, remap = scenes.batch_workload()
    ctx = scenes.to_context(d)
    ctx.render_batch(items, remap, mode=rt.BATCH_MEGAKERNEL)
    ref = _films(ctx, items)
    ref_counts = ctx.batch_counters()
    assert ref_counts["paths"] == sum(it["width"] * it["height"] * it["spp"] for it in items)
    for mode, name in ((rt.BATCH_WAVEFRONT, "wavefront"), (rt.BATCH_PERSISTENT, "persistent")):
        secs = ctx.render_batch(items, remap, mode=mode)
        counts = ctx.batch_counters()
        _say("%s: %.3f s, counters %s (megakernel %s)" % (name, secs, counts, ref_counts))
        for i, it in enumerate(items):
            a = ctx.batch_film(i, it["width"], it["height"], sums=True)
            assert np.array_equal(a[..., 3], ref[i][..., 3]), (name, i)
            np.testing.assert_allclose(a[..., :3], ref[i][..., :3], rtol=1e-9, atol=1e-12, err_msg="%s %d" % (name, i))
        assert counts == ref_counts, (name, counts, ref_counts)
