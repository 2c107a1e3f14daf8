"""Accuracy of the shared FP32 transcendentals (csrc/orf_math.h) against
double-precision libm (numpy / scipy), in ulps of the float result.

The product's kernels and the FP32 oracle both use these functions, so the
comparison to an independent double-precision reference is what pins them.
Sweeps are dense over the ranges the render path uses (angles in [-2pi, 2pi],
cosines in [-1, 1], slopes, log/exp over the RNG's range) plus special values.
A mistyped coefficient or a wrong quadrant rule shows up as an error of many ulps.
"""
import ctypes

import numpy as np
import pytest
from scipy import special

import oracle

FN = {"sin": 0, "cos": 1, "tan": 2, "asin": 3, "acos": 4, "atan": 5, "atan2": 6, "log": 7, "exp": 8, "erf": 9}


def _lib():
    L = oracle.lib()
    L.orc_fmath_eval.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return L


def _eval(name, x, y=None):
    L = _lib()
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.float32)
    out = np.empty_like(x)
    L.orc_fmath_eval(FN[name], len(x), ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(y.ctypes.data),
                     ctypes.c_void_p(out.ctypes.data))
    return out


def _ulp_err(got, want):
    """|got - want| in units of the float spacing at want (subnormal-safe)."""
    w32 = np.abs(want).astype(np.float32)
    sp = np.spacing(np.maximum(w32, np.float32(1.1754944e-38))).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / sp


def _sweep(lo, hi, n=1 << 20):
    x = np.linspace(lo, hi, n).astype(np.float32)
    rng = np.random.default_rng(0)
    return np.concatenate([x, rng.uniform(lo, hi, n).astype(np.float32)])


# name, argument sweep, double reference, max ulp
CASES = [
    ("sin", _sweep(-2 * np.pi, 2 * np.pi), np.sin, 2.0),
    ("cos", _sweep(-2 * np.pi, 2 * np.pi), np.cos, 2.0),
    ("sin", _sweep(-100.0, 100.0, 1 << 18), np.sin, 2.0),
    ("cos", _sweep(-100.0, 100.0, 1 << 18), np.cos, 2.0),
    ("asin", _sweep(-1.0, 1.0), np.arcsin, 2.5),  # pi/2 - 2 asin(..) cancels just above 0.5
    ("acos", _sweep(-1.0, 1.0), np.arccos, 2.0),
    ("atan", _sweep(-50.0, 50.0), np.arctan, 2.5),  # pi/4 + atan(..) cancels just above tan(pi/8)
    ("log", _sweep(1e-7, 1.0), np.log, 2.0),
    ("log", _sweep(1.0, 1e6, 1 << 18), np.log, 2.0),
    ("exp", _sweep(-87.0, 88.0), np.exp, 2.0),
    ("erf", _sweep(-4.5, 4.5), special.erf, 3.0),
]


@pytest.mark.parametrize("name,x,ref,bound", CASES, ids=[c[0] + "_%d" % i for i, c in enumerate(CASES)])
def test_accuracy_against_double_libm(name, x, ref, bound):
    got = _eval(name, x)
    want = ref(x.astype(np.float64))
    err = _ulp_err(got, want)
    # near a zero of sin/cos/log the ulp of the result shrinks; also bound the absolute error
    absok = np.abs(got.astype(np.float64) - want) <= 1e-7 * np.maximum(1.0, np.abs(x.astype(np.float64)))
    ok = (err <= bound) | (absok & (np.abs(want) < 1e-3))
    assert ok.all(), (name, float(err[~ok].max()), x[~ok][:4], got[~ok][:4], want[~ok][:4])


def test_tan_relative():
    x = _sweep(-1.5, 1.5)
    got = _eval("tan", x).astype(np.float64)
    want = np.tan(x.astype(np.float64))
    assert np.all(np.abs(got - want) <= 4 * np.spacing(np.abs(want).astype(np.float32)) + 1e-12)


def test_atan2_quadrants_and_accuracy():
    rng = np.random.default_rng(1)
    y = rng.uniform(-3, 3, 1 << 20).astype(np.float32)
    x = rng.uniform(-3, 3, 1 << 20).astype(np.float32)
    got = _eval("atan2", y, x)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert _ulp_err(got, want).max() <= 2.75
    # special values: zeros of both signs and the axes
    sy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 0.0, -0.0, 1.0, -1.0], np.float32)
    sx = np.array([0.0, 0.0, -0.0, -0.0, 0.0, 0.0, -2.0, -2.0, -1e-30, -1e-30], np.float32)
    got = _eval("atan2", sy, sx)
    want = np.arctan2(sy, sx).astype(np.float32)
    np.testing.assert_allclose(got, want, rtol=2e-7)
    assert np.array_equal(np.signbit(got), np.signbit(want))


def test_special_values():
    nan, inf = np.float32(np.nan), np.float32(np.inf)
    assert np.isnan(_eval("acos", [1.0000001, -1.0000001, nan])).all()
    assert np.isnan(_eval("asin", [1.0000001])).all()
    assert _eval("acos", [1.0])[0] == 0.0
    assert abs(_eval("acos", [-1.0])[0] - np.float32(np.pi)) == 0.0
    lg = _eval("log", [0.0, -1.0, inf, 1.0, 1e-45])
    assert lg[0] == -inf and np.isnan(lg[1]) and lg[2] == inf and lg[3] == 0.0
    assert abs(lg[4] - np.log(1.401298464e-45)) <= 1e-5 * abs(np.log(1.401298464e-45))
    ex = _eval("exp", [0.0, 89.0, -104.0, -100.0, nan])
    assert ex[0] == 1.0 and ex[1] == inf and ex[2] == 0.0 and np.isnan(ex[4])
    assert abs(ex[3] - np.exp(-100.0)) <= 2 * 1.401298464e-45  # subnormal result
    s = _eval("sin", [-0.0, 0.0, inf, nan])
    assert np.signbit(s[0]) and not np.signbit(s[1]) and np.isnan(s[2]) and np.isnan(s[3])
    assert np.array_equal(_eval("erf", [0.0, 4.0, -4.0, 10.0]), np.array([0.0, 1.0, -1.0, 1.0], np.float32))
