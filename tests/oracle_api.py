"""numpy-facing wrappers of the CPU oracle's C entry points (tests only)."""
import ctypes

import numpy as np

import oracle

c_vp = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(c_vp)


def _d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


class Oracle:
    def __init__(self, fp64=False):
        self.lib = oracle.lib(fp64=fp64)
        self.fp64 = fp64
        L = self.lib
        L.orc_tea_key.restype = ctypes.c_uint32
        L.orc_tea_key.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_tea16.restype = ctypes.c_uint32
        L.orc_tea16.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.orc_pcg32_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, c_vp]
        L.orc_tea_stream.argtypes = [ctypes.c_uint32, ctypes.c_int, c_vp, c_vp]
        L.orc_lcg_stream.argtypes = [ctypes.c_uint32, ctypes.c_int, c_vp, c_vp]
        L.orc_md5.argtypes = [ctypes.c_char_p, ctypes.c_uint32, c_vp]
        L.orc_accurate_stream.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, c_vp, c_vp]
        L.orc_gen_bsdf_queries.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64] + [c_vp] * 6
        L.orc_bsdf_ggx_conductor_batch.argtypes = (
            [ctypes.c_int64] + [c_vp] * 6 + [c_vp, c_vp] + [c_vp] * 7 + [ctypes.c_int]
        )
        L.orc_conductor_eta.argtypes = [c_vp] * 4
        L.orc_ggx_G1_batch.argtypes = [ctypes.c_int] + [c_vp] * 6
        L.orc_ggx_ndf_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_double, ctypes.c_double, c_vp]
        L.orc_ggx_vndf_batch.argtypes = [ctypes.c_int, c_vp, c_vp, ctypes.c_double, ctypes.c_double, c_vp]
        L.orc_ggx_sample_vndf_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_float, ctypes.c_float, c_vp, c_vp, c_vp]
        L.orc_fresnel_conductor_batch.argtypes = [ctypes.c_int, c_vp, c_vp, c_vp, c_vp]
        L.orc_fresnel_dielectric_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_float, c_vp]
        L.orc_cosine_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_uint32, c_vp]
        L.orc_beckmann_ndf_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_double, ctypes.c_double, c_vp]
        L.orc_beckmann_vndf_batch.argtypes = [ctypes.c_int, c_vp, c_vp, ctypes.c_double, ctypes.c_double, c_vp]
        L.orc_beckmann_proj_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_double, ctypes.c_double, c_vp]
        L.orc_beckmann_G1_cos_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_double, c_vp]
        L.orc_beckmann_sample_vndf_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_float, ctypes.c_float, c_vp, c_vp,
                                                     c_vp]
        L.orc_erfinv_batch.argtypes = [ctypes.c_int, c_vp, c_vp]
        L.orc_ggx_ms_conductor_batch.argtypes = [ctypes.c_int, c_vp, c_vp, ctypes.c_float, ctypes.c_int,
                                                 ctypes.c_uint32, ctypes.c_int, c_vp, c_vp, c_vp, c_vp]
        L.orc_beckmann_ms_sample_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_float, ctypes.c_float,
                                                   ctypes.c_uint32, ctypes.c_int, c_vp, c_vp, c_vp]
        L.orc_ggx_ms_sample_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                              ctypes.c_int, c_vp, c_vp, c_vp]
        L.orc_ggx_ms_eval_batch.argtypes = [ctypes.c_int, c_vp, c_vp, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                            ctypes.c_int, c_vp]
        L.orc_beckmann_ms_eval_batch.argtypes = L.orc_ggx_ms_eval_batch.argtypes
        L.orc_ms_diffuse_sample_batch.argtypes = [ctypes.c_int, c_vp, ctypes.c_float, ctypes.c_int, c_vp,
                                                  ctypes.c_uint32, ctypes.c_int, c_vp, c_vp, c_vp, c_vp]
        L.orc_ms_diffuse_eval_batch.argtypes = [ctypes.c_int, c_vp, c_vp, ctypes.c_float, ctypes.c_int, c_vp,
                                                ctypes.c_uint32, ctypes.c_int, c_vp]
        L.orc_ms_conductor_eval_batch.argtypes = [ctypes.c_int, c_vp, c_vp, c_vp, ctypes.c_float, ctypes.c_int,
                                                  ctypes.c_uint32, ctypes.c_int, c_vp]
        L.orc_walter_sample_batch.argtypes = [ctypes.c_int, c_vp, c_vp, ctypes.c_float, ctypes.c_int, c_vp]
        L.orc_reflect_weight_batch.argtypes = [ctypes.c_int, c_vp, c_vp, c_vp, ctypes.c_float, ctypes.c_int,
                                               ctypes.c_int, c_vp]

    # ---- RNG
    def tea(self, v0, v1, rounds=16, key=None):
        if key is None:
            key = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)
        k = (ctypes.c_uint32 * 4)(*key)
        return self.lib.orc_tea_key(v0, v1, rounds, k)

    def pcg32(self, seed, seq, n):
        out = np.zeros(n, np.uint32)
        self.lib.orc_pcg32_stream(seed, seq, n, _p(out))
        return out

    def tea_stream(self, seed, n):
        st = np.zeros(n, np.uint32)
        u = np.zeros(n, np.float32)
        self.lib.orc_tea_stream(seed, n, _p(st), _p(u))
        return st, u

    def md5(self, msg):
        out = np.zeros(4, np.uint32)
        self.lib.orc_md5(bytes(msg), len(msg), _p(out))
        return out

    def accurate_stream(self, x, y, n):
        st = np.zeros((n, 2), np.uint32)
        u = np.zeros(n, np.float32)
        self.lib.orc_accurate_stream(x, y, n, _p(st), _p(u))
        return st, u

    def lcg_stream(self, seed, n):
        st = np.zeros(n, np.uint32)
        u = np.zeros(n, np.float32)
        self.lib.orc_lcg_stream(seed, n, _p(st), _p(u))
        return st, u

    # ---- operator workload
    def gen_bsdf_queries(self, n, seed=42, seq=54):
        arrs = [np.zeros(n, np.float32) for _ in range(6)]
        self.lib.orc_gen_bsdf_queries(n, seed, seq, *[_p(a) for a in arrs])
        return arrs

    def conductor_eta(self, ior_re, ior_im):
        er = np.zeros(3)
        ei = np.zeros(3)
        self.lib.orc_conductor_eta(_p(np.asarray(ior_re, np.float64)), _p(np.asarray(ior_im, np.float64)), _p(er), _p(ei))
        return er + 1j * ei

    def bsdf_ggx_conductor(self, wx, wy, wz, alpha, u1, u2, ior_re, ior_im, threads=8):
        n = len(wx)
        ins = [np.ascontiguousarray(a, np.float32) for a in (wx, wy, wz, alpha, u1, u2)]
        outs = [np.zeros(n, np.float32) for _ in range(7)]
        self.lib.orc_bsdf_ggx_conductor_batch(
            n, *[_p(a) for a in ins], _d3(ior_re), _d3(ior_im), *[_p(a) for a in outs], threads
        )
        return outs

    # ---- single functions
    def ggx_G1(self, v, m, nrm, ax, ay):
        n = len(v)
        out = np.zeros(n, np.float32)
        args = [np.ascontiguousarray(a, np.float32) for a in (v, m, nrm)]
        axa = np.full(n, ax, np.float32) if np.isscalar(ax) else np.ascontiguousarray(ax, np.float32)
        aya = np.full(n, ay, np.float32) if np.isscalar(ay) else np.ascontiguousarray(ay, np.float32)
        self.lib.orc_ggx_G1_batch(n, *[_p(a) for a in args], _p(axa), _p(aya), _p(out))
        return out

    def ggx_ndf(self, wm, ax, ay):
        wm = np.ascontiguousarray(wm, np.float64)
        out = np.zeros(len(wm))
        self.lib.orc_ggx_ndf_batch(len(wm), _p(wm), ax, ay, _p(out))
        return out

    def ggx_vndf(self, wi, wm, ax, ay):
        wm = np.ascontiguousarray(wm, np.float64)
        wi = np.ascontiguousarray(wi, np.float64)
        out = np.zeros(len(wm))
        self.lib.orc_ggx_vndf_batch(len(wm), _p(wi), _p(wm), ax, ay, _p(out))
        return out

    def ggx_sample_vndf(self, wi, ax, ay, z1, z2):
        wi = np.ascontiguousarray(wi, np.float32)
        z1 = np.ascontiguousarray(z1, np.float32)
        z2 = np.ascontiguousarray(z2, np.float32)
        out = np.zeros((len(z1), 3), np.float32)
        self.lib.orc_ggx_sample_vndf_batch(len(z1), _p(wi), ax, ay, _p(z1), _p(z2), _p(out))
        return out

    def fresnel_conductor(self, cos_i, ior_re, ior_im):
        c = np.ascontiguousarray(cos_i, np.float32)
        out = np.zeros((len(c), 3), np.float32)
        self.lib.orc_fresnel_conductor_batch(len(c), _p(c), _d3(ior_re), _d3(ior_im), _p(out))
        return out

    def fresnel_dielectric(self, cos_i, eta):
        c = np.ascontiguousarray(cos_i, np.float32)
        out = np.zeros(len(c), np.float32)
        self.lib.orc_fresnel_dielectric_batch(len(c), _p(c), eta, _p(out))
        return out

    def sample_cosine(self, nrm, seed, n):
        nn = np.ascontiguousarray(nrm, np.float32)
        out = np.zeros((n, 3), np.float32)
        self.lib.orc_cosine_batch(n, _p(nn), seed, _p(out))
        return out

    # ---- Beckmann
    def beckmann_ndf(self, wm, ax, ay):
        wm = np.ascontiguousarray(wm, np.float64)
        out = np.zeros(len(wm))
        self.lib.orc_beckmann_ndf_batch(len(wm), _p(wm), ax, ay, _p(out))
        return out

    def beckmann_vndf(self, wi, wm, ax, ay):
        wi = np.ascontiguousarray(wi, np.float64)
        wm = np.ascontiguousarray(wm, np.float64)
        out = np.zeros(len(wm))
        self.lib.orc_beckmann_vndf_batch(len(wm), _p(wi), _p(wm), ax, ay, _p(out))
        return out

    def beckmann_proj(self, wi, ax, ay):
        wi = np.ascontiguousarray(np.atleast_2d(wi), np.float64)
        out = np.zeros(len(wi))
        self.lib.orc_beckmann_proj_batch(len(wi), _p(wi), ax, ay, _p(out))
        return out

    def beckmann_G1_cos(self, cos_t, alpha):
        c = np.ascontiguousarray(cos_t, np.float64)
        out = np.zeros(len(c))
        self.lib.orc_beckmann_G1_cos_batch(len(c), _p(c), alpha, _p(out))
        return out

    def beckmann_sample_vndf(self, wi, ax, ay, z1, z2):
        wi = np.ascontiguousarray(wi, np.float32)
        z1 = np.ascontiguousarray(z1, np.float32)
        z2 = np.ascontiguousarray(z2, np.float32)
        out = np.zeros((len(z1), 3), np.float32)
        self.lib.orc_beckmann_sample_vndf_batch(len(z1), _p(wi), ax, ay, _p(z1), _p(z2), _p(out))
        return out

    def erfinv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(len(x))
        self.lib.orc_erfinv_batch(len(x), _p(x), _p(out))
        return out

    def ggx_ms_sample(self, wi, eta, alpha, n, seed=1, cap=10):
        """(exit directions (n, 3), scattering orders, completed flags) of n walks."""
        wi = np.ascontiguousarray(wi, np.float32)
        wo = np.zeros((n, 3), np.float32)
        order = np.zeros(n, np.int32)
        ok = np.zeros(n, np.int32)
        self.lib.orc_ggx_ms_sample_batch(n, _p(wi), eta, alpha, seed, cap, _p(wo), _p(order), _p(ok))
        return wo, order, ok

    def beckmann_ms_sample(self, wi, eta, alpha, n, seed=1, cap=10):
        """Beckmann dielectric walks: (exit directions (n, 3), orders, completed flags)."""
        wi = np.ascontiguousarray(wi, np.float32)
        wo = np.zeros((n, 3), np.float32)
        order = np.zeros(n, np.int32)
        ok = np.zeros(n, np.int32)
        self.lib.orc_beckmann_ms_sample_batch(n, _p(wi), eta, alpha, seed, cap, _p(wo), _p(order), _p(ok))
        return wo, order, ok

    def ggx_ms_conductor(self, wi, ior, alpha, n, seed=1, cap=10, dist=1):
        """Conductor walks: (exit directions (n, 3), weights (n, 3), orders, completed flags).
        ior: ((re, im),) * 3."""
        wi = np.ascontiguousarray(wi, np.float32)
        c = np.ascontiguousarray(np.asarray(ior, np.float64).ravel())
        wo = np.zeros((n, 3), np.float32)
        w = np.zeros((n, 3), np.float32)
        order = np.zeros(n, np.int32)
        ok = np.zeros(n, np.int32)
        self.lib.orc_ggx_ms_conductor_batch(n, _p(wi), _p(c), alpha, dist, seed, cap, _p(wo), _p(w), _p(order),
                                            _p(ok))
        return wo, w, order, ok

    def ggx_ms_eval(self, wi, wo, eta, alpha, seed=1, cap=10):
        """One multiple-scattering BSDF x cos estimate per row of wo (local frame)."""
        wi = np.ascontiguousarray(wi, np.float32)
        wo = np.ascontiguousarray(wo, np.float32)
        out = np.zeros(len(wo), np.float32)
        self.lib.orc_ggx_ms_eval_batch(len(wo), _p(wi), _p(wo), eta, alpha, seed, cap, _p(out))
        return out

    def beckmann_ms_eval(self, wi, wo, eta, alpha, seed=1, cap=10):
        """ggx_ms_eval on the Beckmann walk."""
        wi = np.ascontiguousarray(wi, np.float32)
        wo = np.ascontiguousarray(wo, np.float32)
        out = np.zeros(len(wo), np.float32)
        self.lib.orc_beckmann_ms_eval_batch(len(wo), _p(wi), _p(wo), eta, alpha, seed, cap, _p(out))
        return out

    def ms_diffuse_sample(self, wi, alpha, rho, n, dist=1, seed=1, cap=100):
        """Rough diffuse walks: (exit directions, last visible normals, weights) each (n, 3), and orders."""
        wi = np.ascontiguousarray(wi, np.float32)
        rho = np.ascontiguousarray(rho, np.float32)
        wo, wm, w = (np.zeros((n, 3), np.float32) for _ in range(3))
        order = np.zeros(n, np.int32)
        self.lib.orc_ms_diffuse_sample_batch(n, _p(wi), alpha, dist, _p(rho), seed, cap, _p(wo), _p(wm), _p(w),
                                             _p(order))
        return wo, wm, w, order

    def ms_diffuse_eval(self, wi, wo, alpha, rho, dist=1, seed=1, cap=100):
        """One rough diffuse walk estimate of BSDF x cos (n, 3) per row of wo."""
        wi = np.ascontiguousarray(wi, np.float32)
        wo = np.ascontiguousarray(wo, np.float32)
        rho = np.ascontiguousarray(rho, np.float32)
        out = np.zeros((len(wo), 3), np.float32)
        self.lib.orc_ms_diffuse_eval_batch(len(wo), _p(wi), _p(wo), alpha, dist, _p(rho), seed, cap, _p(out))
        return out

    def ms_conductor_eval(self, wi, wo, ior, alpha, dist=1, seed=1, cap=10):
        """One conductor walk estimate of BSDF x cos (n, 3) per row of wo; ior ((re, im),) * 3."""
        wi = np.ascontiguousarray(wi, np.float32)
        wo = np.ascontiguousarray(wo, np.float32)
        c = np.ascontiguousarray(np.asarray(ior, np.float64).ravel())
        out = np.zeros((len(wo), 3), np.float32)
        self.lib.orc_ms_conductor_eval_batch(len(wo), _p(wi), _p(wo), _p(c), alpha, dist, seed, cap, _p(out))
        return out

    def walter_sample(self, z1, z2, alpha, dist):
        z1 = np.ascontiguousarray(z1, np.float32)
        z2 = np.ascontiguousarray(z2, np.float32)
        m = np.zeros((len(z1), 3), np.float32)
        self.lib.orc_walter_sample_batch(len(z1), _p(z1), _p(z2), alpha, dist, _p(m))
        return m

    def reflect_weight(self, wi, z1, z2, alpha, dist, model):
        wi = np.ascontiguousarray(wi, np.float32)
        z1 = np.ascontiguousarray(z1, np.float32)
        z2 = np.ascontiguousarray(z2, np.float32)
        out = np.zeros(len(z1), np.float32)
        self.lib.orc_reflect_weight_batch(len(z1), _p(wi), _p(z1), _p(z2), alpha, dist, model, _p(out))
        return out
