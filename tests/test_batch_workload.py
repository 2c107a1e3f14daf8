"""CPU checks of the batched workload definition and the batch C-ABI's argument checks."""
import numpy as np
import pytest

from open_rendering_framework_amd import scenes


def test_pcg32_matches_published_demo_vector():
    g = scenes.Pcg32(42, 54)
    assert [g.next() for _ in range(6)] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_batch_workload_properties():
    d, items, remap = scenes.batch_workload(level=1, env_size=(64, 32))
    assert len(items) == 64 and remap.shape == (64, 5)
    for it in items:
        assert 64 <= it["width"] <= 1024 and it["width"] % 8 == 0
        assert 64 <= it["height"] <= 1024 and it["height"] % 8 == 0
        assert it["spp"] in (4, 16, 64, 256)
        assert it["material"] in ("lambert", "conductor", "dielectric")
    assert set(remap[:, 1]) <= {1, 3, 4}
    assert {it["material"] for it in items} == {"lambert", "conductor", "dielectric"}
    # deterministic per seed, different across seeds
    _, items2, _ = scenes.batch_workload(level=1, env_size=(64, 32))
    assert items == items2
    _, items3, _ = scenes.batch_workload(seed=99, level=1, env_size=(64, 32))
    assert items != items3


def test_batch_rejects_bad_remap():
    from open_rendering_framework_amd import rt

    d, items, remap = scenes.batch_workload(n=2, level=1, env_size=(64, 32), spp_scale=1 / 64, size_scale=0.1)
    bad = remap.copy()
    bad[0, 1] = 99
    ctx = rt.Context.__new__(rt.Context)
    ctx.handle = None  # no device here: only the argument checks are reachable
    with pytest.raises(rt.OrfError):
        rt.call("orf_render_batch", None, ctx._items(items), 2, None, 0, 0, None, None, None)
