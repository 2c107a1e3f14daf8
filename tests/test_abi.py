"""CPU checks of the C-ABI library: it loads, exports every declared symbol,
reports errors through status codes, and its host-side workload generator
reproduces the oracle's stream bit for bit.  No GPU compute is issued."""
import ctypes
import os
import re

import numpy as np
import pytest

from open_rendering_framework_amd import INCLUDE_DIR


def _declared_symbols():
    syms = set()
    for name in os.listdir(INCLUDE_DIR):
        if name.endswith(".h"):
            text = open(os.path.join(INCLUDE_DIR, name)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms.update(re.findall(r"\b(orf_[a-z0-9_]+)\s*\(", text))
    return sorted(syms)


def test_library_exports_every_declared_symbol():
    from open_rendering_framework_amd import rt

    syms = _declared_symbols()
    assert len(syms) >= 4
    missing = [s for s in syms if not hasattr(rt.lib, s)]
    assert not missing, missing


def test_version_and_error_path():
    from open_rendering_framework_amd import rt

    assert "gfx950" in rt.version()
    with pytest.raises(rt.OrfError) as ei:
        rt.call("orf_bsdf_ggx_conductor_batch", -1, *([None] * 6), None, None, *([None] * 7), None)
    assert ei.value.status == 1
    assert "n < 0" in str(ei.value)
    # n == 0 is a valid empty batch (no launch)
    rt.call("orf_bsdf_ggx_conductor_batch", 0, *([None] * 6), rt._dbl3([1, 1, 1]), rt._dbl3([0, 0, 0]),
            *([None] * 7), None)


def test_generator_matches_oracle_bitwise(orc):
    from open_rendering_framework_amd import rt

    n = 50000
    got = [t.numpy() for t in rt.gen_bsdf_queries(n, 42, 54)]
    want = orc.gen_bsdf_queries(n, 42, 54)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)
    wx, wy, wz, al, u1, u2 = want
    assert (wz >= 0).all() and (wz < 1).all()
    np.testing.assert_allclose(wx * wx + wy * wy + wz * wz, 1.0, atol=1e-6)
    assert al.min() >= 0.05 and al.max() <= 1.0
