"""MI355X-native render runtime for the open_rendering_framework render path.

The compute path is liborf_rt.so (hand-written gfx950 HIP kernels behind a
C-ABI declared in include/orf_rt.h).  This package loads it with ctypes and
offers a thin Python view of the same host API; torch supplies device memory
and streams only.  There is no CPU fallback: if the library is missing the
import of `open_rendering_framework_amd.rt` raises.
"""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PACKAGE_DIR, "liborf_rt.so")
REPO_ROOT = os.path.dirname(PACKAGE_DIR)
INCLUDE_DIR = os.path.join(REPO_ROOT, "include")

__all__ = ["PACKAGE_DIR", "LIB_PATH", "REPO_ROOT", "INCLUDE_DIR", "build"]


def build(jobs=8):
    """Compile liborf_rt.so for gfx950 in-tree (make in csrc/)."""
    import subprocess

    subprocess.run(["make", "-s", "-j%d" % jobs, "-C", os.path.join(PACKAGE_DIR, "csrc")], check=True)
