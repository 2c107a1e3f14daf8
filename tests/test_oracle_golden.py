"""The oracle reproduces its committed golden fixtures exactly (regression pin)."""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_bsdf_fixture_reproduced(orc):
    f = np.load(os.path.join(HERE, "bsdf_ggx_conductor_4096.npz"))
    ins = [f[k] for k in ("wo_x", "wo_y", "wo_z", "alpha", "u1", "u2")]
    got = np.stack(orc.bsdf_ggx_conductor(*ins, f["ior_re"], f["ior_im"]))
    np.testing.assert_array_equal(got, f["out"])
    got_c = np.stack(orc.bsdf_ggx_conductor(*ins, f["ior_re_c"], f["ior_im_c"]))
    np.testing.assert_array_equal(got_c, f["out_c"])
