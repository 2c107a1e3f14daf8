"""Writes the committed golden fixtures from the CPU oracle.

The reference ships no golden vectors for its render path, so these fixtures
are the oracle's own outputs on seeded inputs (inputs and expected outputs),
kept small.  They pin the oracle against regressions and give the GPU parity
tests a fixed target.  Run from the repository root:
    python -m tests.golden.make_golden
"""
import os

import numpy as np

from tests.oracle_api import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
IOR_RE = [0.27, 0.27, 0.27]
IOR_IM = [2.78, 2.78, 2.78]


def bsdf_fixture(n=4096):
    orc = Oracle()
    wx, wy, wz, al, u1, u2 = orc.gen_bsdf_queries(n, 42, 54)
    # a second, coloured IOR exercises the per-channel path
    outs = orc.bsdf_ggx_conductor(wx, wy, wz, al, u1, u2, IOR_RE, IOR_IM)
    outs_c = orc.bsdf_ggx_conductor(wx, wy, wz, al, u1, u2, [0.2, 1.1, 1.5], [3.9, 2.6, 1.9])
    np.savez_compressed(os.path.join(HERE, "bsdf_ggx_conductor_4096.npz"),
             wo_x=wx, wo_y=wy, wo_z=wz, alpha=al, u1=u1, u2=u2,
             ior_re=np.array(IOR_RE), ior_im=np.array(IOR_IM),
             out=np.stack(outs), ior_re_c=np.array([0.2, 1.1, 1.5]), ior_im_c=np.array([3.9, 2.6, 1.9]),
             out_c=np.stack(outs_c))


if __name__ == "__main__":
    bsdf_fixture()
