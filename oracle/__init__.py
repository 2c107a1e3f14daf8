"""ctypes loader for the CPU oracle (test infrastructure only).

The oracle is a plain-C restatement of the reference render path (see
orf_oracle.h).  Only tests, the smoke check and the benchmark's CPU-baseline
leg load it; the product package never imports this module.
"""
import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")


def build():
    """Compile both oracle variants with gcc (make in this directory)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(name):
    path = os.path.join(_BUILD, name)
    if not os.path.exists(path):
        build()
    return ctypes.CDLL(path)


_libs = {}


def lib(fp64=False):
    """Return the FP32 (default) or FP64 oracle shared library."""
    key = "fp64" if fp64 else "fp32"
    if key not in _libs:
        _libs[key] = _load("liboracle_fp64.so" if fp64 else "liboracle.so")
    return _libs[key]
