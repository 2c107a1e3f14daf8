"""ctypes view of liborf_rt.so (the C-ABI in include/orf_rt.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (one runtime per process).  Every call checks the returned
status and raises OrfError with orf_last_error()'s text.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the library load: shared HIP runtime)

from . import LIB_PATH

c_f32p = ctypes.POINTER(ctypes.c_float)
c_vp = ctypes.c_void_p
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_u32 = ctypes.c_uint32
c_i32 = ctypes.c_int32
c_dblp = ctypes.POINTER(ctypes.c_double)


class OrfError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero orf_status."""

    def __init__(self, status, msg):
        super().__init__("orf_status %d: %s" % (status, msg))
        self.status = status


if not os.path.exists(LIB_PATH):
    raise ImportError(
        "liborf_rt.so not built at %s; run open_rendering_framework_amd.build() "
        "(no CPU fallback exists)" % LIB_PATH
    )

lib = ctypes.CDLL(LIB_PATH)
lib.orf_last_error.restype = ctypes.c_char_p
lib.orf_version.restype = ctypes.c_char_p

# name -> argtypes; restype is orf_status (int) unless listed in _SPECIAL
_SIGS = {
    "orf_bsdf_ggx_conductor_batch": [c_i64] + [c_vp] * 6 + [c_dblp, c_dblp] + [c_vp] * 7 + [c_vp],
    "orf_gen_bsdf_queries": [c_i64, c_u64, c_u64] + [c_vp] * 6,
}
for _name, _args in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = ctypes.c_int


def check(status):
    if status != 0:
        raise OrfError(status, lib.orf_last_error().decode())


def call(name, *args):
    """Call a C-ABI entry point and raise OrfError on failure."""
    check(getattr(lib, name)(*args))


def ptr(t):
    """Device/host address of a contiguous torch tensor (None -> NULL)."""
    if t is None:
        return None
    assert t.is_contiguous(), "tensor must be contiguous"
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def version():
    return lib.orf_version().decode()


def _dbl3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def gen_bsdf_queries(n, seed=42, seq=54):
    """Synthetic operator workload (host float32 tensors wo_x, wo_y, wo_z, alpha, u1, u2)."""
    outs = [torch.empty(n, dtype=torch.float32) for _ in range(6)]
    call("orf_gen_bsdf_queries", n, seed, seq, *[ptr(t) for t in outs])
    return outs


def bsdf_ggx_conductor_batch(wo_x, wo_y, wo_z, alpha, u1, u2, ior_re, ior_im, out=None, stream=None):
    """Rough-conductor VNDF sampling step over SoA device tensors.

    Returns (wi_x, wi_y, wi_z, pdf, w_r, w_g, w_b) as float32 device tensors.
    """
    n = wo_x.numel()
    for t in (wo_x, wo_y, wo_z, alpha, u1, u2):
        if t.dtype != torch.float32 or not t.is_cuda or t.numel() != n:
            raise ValueError("inputs must be float32 CUDA tensors of equal length")
    if out is None:
        out = [torch.empty(n, dtype=torch.float32, device=wo_x.device) for _ in range(7)]
    call(
        "orf_bsdf_ggx_conductor_batch",
        n,
        *[ptr(t) for t in (wo_x, wo_y, wo_z, alpha, u1, u2)],
        _dbl3(ior_re),
        _dbl3(ior_im),
        *[ptr(t) for t in out],
        stream_ptr(stream),
    )
    return out
