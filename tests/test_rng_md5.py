"""The translucent programs' 64-bit random stream (rnd_accurate, MD5 of ORF/md5.h).

Pins, each independent of the code under test:
- MD5 itself: the RFC 1321 appendix A.5 test suite (the messages of at most 55
  bytes, which fit one block as the reference's hash requires) and Python's
  hashlib on random messages of every length 0..55;
- a draw: the next state is the first two digest words of the state's 8 bytes
  (hashlib), the value is float32(float64(low 63 bits) / 2^63) (numpy's
  conversions round to nearest even, as the reference's casts do);
- the double rounding of that value on crafted integers where rounding once
  would give another float.
The oracle (C restatement) and the product (the kernels' own code, run on the
host through orf_rng_accurate_stream / orf_rng_unit63) are held to all of them.
"""
import ctypes
import hashlib
import struct

import numpy as np
import pytest

RFC1321 = [
    (b"", "d41d8cd98f00b204e9800998ecf8427e"),
    (b"a", "0cc175b9c0f1b6a831c399e269772661"),
    (b"abc", "900150983cd24fb0d6963f7d28e17f72"),
    (b"message digest", "f96b697d7cb7938d525a2f31aaf161d0"),
    (b"abcdefghijklmnopqrstuvwxyz", "c3fcd3d76192e4007dfb496cca67e13b"),
]


def _words(digest):
    return [int(x) for x in np.frombuffer(digest, "<u4")]


def _expected_stream(x, y, n):
    states, vals = [], []
    for _ in range(n):
        h = _words(hashlib.md5(struct.pack("<II", x, y)).digest())
        x, y = h[0], h[1]
        low63 = ((y << 32) | x) & ((1 << 63) - 1)
        states.append((x, y))
        vals.append(np.float32(np.float64(np.uint64(low63)) / np.float64(2.0**63)))
    return np.array(states, np.uint32), np.array(vals, np.float32)


@pytest.fixture(scope="module")
def product():
    from open_rendering_framework_amd import rt

    L = rt.lib
    L.orf_rng_accurate_stream.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p]
    L.orf_rng_unit63.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]

    class P:
        @staticmethod
        def stream(x, y, n):
            st = np.zeros((n, 2), np.uint32)
            u = np.zeros(n, np.float32)
            rt.check(L.orf_rng_accurate_stream(x, y, n, st.ctypes.data, u.ctypes.data))
            return st, u

        @staticmethod
        def unit63(ls):
            a = np.ascontiguousarray(ls, np.uint64)
            out = np.zeros(len(a), np.float32)
            rt.check(L.orf_rng_unit63(a.ctypes.data, len(a), out.ctypes.data))
            return out

    return P


@pytest.mark.parametrize("msg,hexd", RFC1321)
def test_oracle_md5_rfc1321_suite(orc, msg, hexd):
    assert _words(bytes.fromhex(hexd)) == [int(v) for v in orc.md5(msg)]


def test_oracle_md5_matches_hashlib_every_length(orc):
    rng = np.random.default_rng(3)
    for n in range(56):
        msg = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert _words(hashlib.md5(msg).digest()) == [int(v) for v in orc.md5(msg)], n


@pytest.mark.parametrize("seed", [(0, 0), (1, 0), (0x9E3779B9, 0x9E3779B9), (0xFFFFFFFF, 0x12345678)])
def test_oracle_accurate_stream(orc, seed):
    want_st, want_u = _expected_stream(*seed, 64)
    st, u = orc.accurate_stream(*seed, 64)
    assert np.array_equal(st, want_st)
    assert np.array_equal(u, want_u)


@pytest.mark.parametrize("seed", [(0, 0), (1, 0), (0x9E3779B9, 0x9E3779B9), (0xFFFFFFFF, 0x12345678)])
def test_product_accurate_stream(product, seed):
    want_st, want_u = _expected_stream(*seed, 64)
    st, u = product.stream(*seed, 64)
    assert np.array_equal(st, want_st)
    assert np.array_equal(u, want_u)


def test_product_stream_from_pixel_seeds(product, orc):
    # both words seeded with the pixel seed, as the camera does
    for pix in range(0, 4096, 97):
        s = orc.tea(pix, 5)
        st, u = product.stream(s, s, 8)
        want_st, want_u = _expected_stream(s, s, 8)
        assert np.array_equal(st, want_st) and np.array_equal(u, want_u)


def _crafted():
    ls = [0, 1, 2, 3, (1 << 63) - 1, (1 << 62), (1 << 52), (1 << 53) + 1, (1 << 24) - 1, (1 << 24) + 1]
    # leading bit 62: the double ulp is 2^10, the float ulp 2^39
    base = 1 << 62
    ls += [base | (1 << 9),                    # tie at the double step, even: down
           base | (1 << 10) | (1 << 9),        # tie at the double step, odd: up
           base | (1 << 38),                   # exact tie at the float step: even, down
           base | (1 << 39) | (1 << 38),       # exact tie at the float step: odd, up
           base | (1 << 38) | (1 << 9),        # tie, tie: down twice (once would round up)
           base | (1 << 38) | (1 << 9) | 1,    # just above the double tie: up, then a float tie
           base | ((1 << 39) - 1)]             # rounds up to the next float
    rng = np.random.default_rng(11)
    ls += [int(v) for v in rng.integers(0, 1 << 63, 2000, dtype=np.uint64)]
    for p in range(25, 63):  # random bits below every leading position
        ls += [(1 << p) | int(v) for v in rng.integers(0, 1 << p, 20, dtype=np.uint64)]
    return ls


def test_product_unit63_double_rounding(product):
    ls = _crafted()
    want = np.array([np.float32(np.float64(np.uint64(v)) / np.float64(2.0**63)) for v in ls], np.float32)
    got = product.unit63(ls)
    assert np.array_equal(got, want)
    # the crafted tie-tie case really needs two roundings
    tt = (1 << 62) | (1 << 38) | (1 << 9)
    assert float(np.float32(tt / 2.0**63)) == 0.5
    assert float(np.float32(np.float64(np.longdouble(tt) / np.longdouble(2.0**63)))) == 0.5
    assert np.float32(np.uint64(tt)) != np.float32(np.uint64(1 << 62))  # one rounding goes up


def test_oracle_accurate_uniform(orc):
    u = np.concatenate([orc.accurate_stream(orc.tea(s, 9), orc.tea(s, 9), 32)[1] for s in range(2000)])
    assert u.min() >= 0.0 and u.max() <= 1.0
    hist, _ = np.histogram(u, bins=50, range=(0, 1))
    expected = len(u) / 50
    assert ((hist - expected) ** 2 / expected).sum() < 100.0
