"""Hands a SceneDesc to the CPU oracle renderer (tests and the CPU baseline only)."""
import ctypes

import numpy as np

import oracle

c_vp = ctypes.c_void_p


class OrcMaterial(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("microfacet_model", ctypes.c_int32),
                ("normal_distribution", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("diffuse", ctypes.c_float * 3), ("radiance", ctypes.c_float * 3),
                ("absorption", ctypes.c_float * 3), ("roughness", ctypes.c_float * 2), ("ior", ctypes.c_float),
                ("ior_re", ctypes.c_double * 3), ("ior_im", ctypes.c_double * 3),
                ("sss_absorption", ctypes.c_float * 3), ("sss_scattering", ctypes.c_float * 3),
                ("sss_meancosine", ctypes.c_float * 3), ("sss_scale", ctypes.c_float), ("dipole_model", ctypes.c_int32),
                ("pad2", ctypes.c_int32), ("flat_flags", ctypes.c_int32), ("flat_threshold", ctypes.c_float)]


class OrcLight(ctypes.Structure):
    _fields_ = [("type", ctypes.c_uint32), ("radiance", ctypes.c_float * 3), ("params", ctypes.c_float * 3),
                ("shape", ctypes.c_float * 3), ("material", ctypes.c_int32)]


class OrcTriLight(ctypes.Structure):
    _fields_ = [("v", ctypes.c_float * 9), ("n", ctypes.c_float * 9), ("has_normals", ctypes.c_uint32),
                ("area", ctypes.c_float)]


class OrcSceneDesc(ctypes.Structure):
    _fields_ = [("nverts", ctypes.c_int64), ("ntris", ctypes.c_int64), ("nnormals", ctypes.c_int64),
                ("verts", c_vp), ("idx", c_vp), ("normals", c_vp), ("nidx", c_vp), ("tri_mat", c_vp),
                ("nmats", ctypes.c_int32), ("mats", c_vp), ("nlights", ctypes.c_int32), ("lights", c_vp),
                ("ntrilights", ctypes.c_int32), ("trilights", c_vp), ("env", c_vp), ("env_w", ctypes.c_int32),
                ("env_h", ctypes.c_int32), ("env_is", ctypes.c_int32), ("eye", ctypes.c_float * 3),
                ("U", ctypes.c_float * 3), ("V", ctypes.c_float * 3), ("W", ctypes.c_float * 3),
                ("max_depth", ctypes.c_int32), ("scene_epsilon", ctypes.c_float)]


_KIND = {"diffuse_shader.cu": 0, "metallic_shader.cu": 1, "arealight_shader.cu": 2, "rough_transparent_shader.cu": 3,
         "lambertian_interface_shader.cu": 4, "subsurface_scattering_shader.cu": 5,
         "rough_subsurface_scattering_shader.cu": 6, "transparent_shader.cu": 7, "normal_shader.cu": 8,
         "flat_shader.cu": 9, "rough_diffuse_shader.cu": 10}
# Wholemilk (ORF/RoughTranslucentMaterial.cpp:258-263) and the default scale
SSS_DEFAULTS = {"absorption": (0.0011, 0.0024, 0.014), "scattering": (2.55, 3.21, 3.77),
                "meancosine": (0.0, 0.0, 0.0), "scale": (100.0,)}


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class OracleScene:
    """An oracle scene built from a SceneDesc; keeps the arrays alive."""

    def __init__(self, desc, fp64=False):
        self.lib = L = oracle.lib(fp64=fp64)
        L.orc_scene_create.restype = c_vp
        L.orc_scene_create.argtypes = [ctypes.POINTER(OrcSceneDesc)]
        L.orc_scene_destroy.argtypes = [c_vp]
        L.orc_render.argtypes = [c_vp] + [ctypes.c_int] * 6 + [ctypes.c_uint32, ctypes.c_uint32, c_vp, c_vp,
                                                               ctypes.c_int]
        L.orc_render_samples.argtypes = [c_vp, ctypes.c_int, ctypes.c_int, c_vp, c_vp, ctypes.c_int,
                                         ctypes.c_uint32, c_vp, ctypes.c_int]
        L.orc_depth_samples.argtypes = [c_vp, ctypes.c_int, ctypes.c_int, c_vp, c_vp, ctypes.c_int, ctypes.c_uint32,
                                        c_vp]
        L.orc_env_tables.argtypes = [c_vp] * 5
        L.orc_env_sample_batch.argtypes = [c_vp, ctypes.c_uint32, ctypes.c_int, c_vp, c_vp, c_vp]
        L.orc_env_bsearch.restype = ctypes.c_int
        L.orc_env_bsearch.argtypes = [c_vp, ctypes.c_int, ctypes.c_double]
        L.orc_closest_hits.argtypes = [c_vp, ctypes.c_int, c_vp, c_vp, ctypes.c_float, c_vp, c_vp, ctypes.c_int]
        L.orc_probes_set.argtypes = [c_vp, ctypes.c_int, c_vp, c_vp, ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.orc_probe_pass.argtypes = [c_vp, ctypes.c_uint32, c_vp, c_vp, c_vp, ctypes.c_int]
        L.orc_probes_values.argtypes = [c_vp, c_vp, c_vp, c_vp]
        L.orc_sss_coefficients.argtypes = [c_vp, ctypes.c_int, c_vp]
        L.orc_dipole_eval.argtypes = [c_vp, ctypes.c_int, ctypes.c_int, c_vp, c_vp]
        L.orc_dirpole_eval.argtypes = [c_vp, ctypes.c_int, ctypes.c_int, c_vp, c_vp]
        L.orc_dirpole_S_infinite_eval.argtypes = [c_vp, ctypes.c_int, ctypes.c_int, c_vp, c_vp]
        L.orc_sss_gather_eval.argtypes = [c_vp, ctypes.c_int, c_vp, c_vp, ctypes.c_int, ctypes.c_uint32, c_vp]
        L.orc_diff_set.argtypes = [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_vp, ctypes.c_int,
                                   ctypes.c_int, c_vp, c_vp, c_vp, c_vp, c_vp]
        L.orc_diff_eval.restype = ctypes.c_double
        L.orc_diff_set_sss.argtypes = [c_vp, ctypes.c_int, c_vp, c_vp]
        L.orc_diff_eval.argtypes = [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_vp, c_vp, ctypes.c_uint32,
                                    ctypes.c_uint32, c_vp, c_vp, ctypes.c_int]
        self.desc = desc
        V, T, N, NI, M = desc.flat()
        self.arrays = [np.ascontiguousarray(V), np.ascontiguousarray(T), np.ascontiguousarray(N),
                       np.ascontiguousarray(NI), np.ascontiguousarray(M)]
        # disk and sphere lights carry an emitter material each, after the scene's
        emitters = [l for l in desc.lights if l.kind in ("disk", "sphere")]
        nm = len(desc.materials)
        mats = (OrcMaterial * max(1, nm + len(emitters)))()
        for k, l in enumerate(emitters):
            mats[nm + k].kind = 2
            mats[nm + k].radiance[:] = [float(x) for x in l.radiance]
        for i, m in enumerate(desc.materials):
            mm = mats[i]
            mm.kind = _KIND[m.program]
            mm.microfacet_model = m.uints.get("microfacet_model", (1,))[0]
            # RoughTranslucentMaterial defaults to Beckmann (ORF/RoughTranslucentMaterial.cpp:17)
            mm.normal_distribution = m.uints.get("normal_distribution", (0,) if mm.kind == 6 else (1,))[0]
            for k, v in enumerate(m.floats.get("diffuse_color", (0.8, 0.8, 0.8))):
                mm.diffuse[k] = v
            for k, v in enumerate(m.floats.get("radiance", (1.0, 1.0, 1.0))):
                mm.radiance[k] = v
            for k, v in enumerate(m.floats.get("glass_absorption", (0.0, 0.0, 0.0))):
                mm.absorption[k] = v
            r = m.floats.get("roughness", (0.1, 0.1))
            mm.roughness[0] = r[0]
            mm.roughness[1] = r[1] if len(r) > 1 else r[0]
            mm.ior = m.floats.get("ior", (1.3,) if mm.kind in (5, 6, 7) else (1.5,))[0]
            for key, field in (("absorption", "sss_absorption"), ("scattering", "sss_scattering"),
                               ("meancosine", "sss_meancosine")):
                getattr(mm, field)[:] = [float(x) for x in m.floats.get(key, SSS_DEFAULTS[key])]
            mm.sss_scale = m.floats.get("scale", SSS_DEFAULTS["scale"])[0]
            mm.dipole_model = m.uints.get("dipole_model", (0,))[0]
            if mm.kind == 10:  # RoughDiffuseMaterial defaults (ORF/RoughDiffuseMaterial.cpp:7-17)
                mm.diffuse[:] = [float(x) for x in m.floats.get("diffuse_color", (1.0, 1.0, 1.0))]
                mm.ior = m.floats.get("ior", (1.3,))[0]
                mm.normal_distribution = m.uints.get("normal_distribution", (0,))[0]
                mm.microfacet_model = m.uints.get("microfacet_model", (0,))[0]
            if mm.kind == 9:  # FlatMaterial defaults (ORF/FlatMaterial.cpp:5-21)
                mm.diffuse[:] = [float(x) for x in m.floats.get("diffuse_color", (1.0, 1.0, 1.0))]
                mm.radiance[:] = [float(x) for x in m.floats.get("highlight_color", (1.0, 1.0, 1.0))]
                r = m.floats.get("roughness", (0.999, 0.999))
                mm.roughness[0], mm.roughness[1] = r[0], r[-1]
                mm.normal_distribution = m.uints.get("normal_distribution", (0,))[0]
                mm.ior = m.floats.get("ior", (1.3,))[0]
                mm.flat_flags = (1 if m.uints.get("flat_shadow", (0,))[0] == 1 else 0) | \
                    (2 if m.uints.get("highlight", (0,))[0] == 1 else 0)
                mm.flat_threshold = m.floats.get("highlight_threshold", (0.0,))[0]
            ior = m.ior_complex or ((1.5, 0.0),) * 3
            for k in range(3):
                mm.ior_re[k] = ior[k][0]
                mm.ior_im[k] = ior[k][1]
        lights, tls = [], []
        n_em = 0
        for l in desc.lights:
            ol = OrcLight()
            ol.radiance[:] = list(l.radiance)
            if l.kind == "point":
                ol.type = 0
                ol.params[:] = list(l.position)
            elif l.kind == "directional":
                ol.type = 1
                d = np.asarray(l.direction, np.float64)
                d = (d / np.linalg.norm(d)).astype(np.float32)
                ol.params[:] = [float(x) for x in d]
            elif l.kind in ("disk", "sphere"):
                ol.type = 3 if l.kind == "disk" else 4
                ol.params[:] = list(l.position)
                ol.shape[:] = [l.radius, l.theta, l.phi]
                ol.material = nm + n_em
                n_em += 1
            else:
                ol.type = 2
                n = len(l.tri_verts)
                ol.params[:] = [float(n), float(len(tls)), float(len(tls) + n - 1)]
                for t in range(n):
                    tl = OrcTriLight()
                    v = np.asarray(l.tri_verts[t], np.float32)
                    tl.v[:] = [float(x) for x in v]
                    if l.tri_normals is not None:
                        tl.n[:] = [float(x) for x in np.asarray(l.tri_normals[t], np.float32)]
                        tl.has_normals = 1
                    a, b, c = v[0:3], v[3:6], v[6:9]
                    tl.area = np.float32(0.5 * float(np.linalg.norm(np.cross(b - c, a - c).astype(np.float32))))
                    tls.append(tl)
            lights.append(ol)
        self.mats = mats
        self.lights = (OrcLight * max(1, len(lights)))(*lights)
        self.tls = (OrcTriLight * max(1, len(tls)))(*tls)
        self.env = None if desc.env is None else np.ascontiguousarray(desc.env, np.float32)
        d = OrcSceneDesc()
        d.nverts, d.ntris, d.nnormals = len(V), len(T), len(N)
        d.verts, d.idx, d.normals, d.nidx, d.tri_mat = [_p(a) for a in self.arrays]
        if len(N) == 0:
            d.normals = None
            d.nidx = None
        d.nmats = nm + len(emitters)
        d.mats = ctypes.cast(self.mats, c_vp)
        d.nlights = len(lights)
        d.lights = ctypes.cast(self.lights, c_vp)
        d.ntrilights = len(tls)
        d.trilights = ctypes.cast(self.tls, c_vp)
        if self.env is not None:
            d.env = _p(self.env)
            d.env_h, d.env_w = self.env.shape[:2]
            d.env_is = 1 if desc.env_importance else 0
        for name in ("eye", "U", "V", "W"):
            getattr(d, name)[:] = list(getattr(desc, name))
        d.max_depth = desc.max_depth
        d.scene_epsilon = desc.scene_epsilon
        self.cdesc = d
        self.handle = L.orc_scene_create(ctypes.byref(d))

    def close(self):
        if self.handle:
            self.lib.orc_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def render(self, width, height, frame0=0, spp=1, rect=None, threads=8):
        x0, y0, x1, y1 = rect if rect is not None else (0, 0, width, height)
        out = np.zeros((y1 - y0, x1 - x0, 4), np.float64)
        cnt = np.zeros(4, np.uint64)
        self.lib.orc_render(self.handle, width, height, x0, y0, x1, y1, frame0, spp, _p(out), _p(cnt), threads)
        return out, dict(zip(("nonfinite", "ext_rays", "shadow_rays", "paths"), [int(x) for x in cnt]))

    def samples(self, width, height, px, py, frame, threads=8):
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        out = np.zeros((len(px), 3), np.float64)
        self.lib.orc_render_samples(self.handle, width, height, _p(px), _p(py), len(px), frame, _p(out), threads)
        return out

    def depth_samples(self, width, height, px, py, frame):
        """Depth integrator: (n, 4) normal colour and hit distance for one frame."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        out = np.zeros((len(px), 4))
        self.lib.orc_depth_samples(self.handle, width, height, _p(px), _p(py), len(px), frame, _p(out))
        return out

    def env_tables(self):
        h, w = self.env.shape[:2]
        out = [np.zeros(h), np.zeros(h), np.zeros(h * w), np.zeros(h * w)]
        self.lib.orc_env_tables(self.handle, *[_p(a) for a in out])
        return out

    def env_samples(self, n, seed=1):
        d = np.zeros((n, 3))
        ps = np.zeros(n)
        pe = np.zeros(n)
        self.lib.orc_env_sample_batch(self.handle, seed, n, _p(d), _p(ps), _p(pe))
        return d, ps, pe

    def closest_hits(self, o, d, tmin=1e-4, brute_force=False):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        t = np.zeros(len(o), np.float32)
        prim = np.zeros(len(o), np.int32)
        self.lib.orc_closest_hits(self.handle, len(o), _p(o), _p(d), tmin, _p(t), _p(prim), 1 if brute_force else 0)
        return t, prim

    # ------------------------------------------------ differentiable form
    def diff_set(self, tex_size, albedo_material, roughness_material, albedo=None, roughness=None,
                 albedo_frozen=None, roughness_frozen=None, albedo_uv=(0, (-10.0, -10.0, 10.0, 10.0)),
                 roughness_uv=(1, (0.0, 0.0, 0.0, 0.0))):
        """Textured albedo (tex_h, tex_w, 3) on a Lambert material and roughness
        (tex_h, tex_w) on a GGX conductor; *_frozen (default: the same values)
        drive sampling and Russian roulette."""
        tw, th = tex_size
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
        self._diff_keep = [f64(albedo), f64(albedo_frozen), f64(roughness), f64(roughness_frozen),
                           np.asarray(albedo_uv[1], np.float64), np.asarray(roughness_uv[1], np.float64)]
        a, af, r, rf, au, ru = self._diff_keep
        self.lib.orc_diff_set(self.handle, tw, th, albedo_material, albedo_uv[0], _p(au), roughness_material,
                              roughness_uv[0], _p(ru), _p(a), _p(af), _p(r), _p(rf))

    def diff_set_sss(self, material, eval6, frozen6=None):
        """Absorption RGB + scattering RGB (1/mm) of a translucent material as
        parameters: shade with eval6, sample with frozen6 (default eval6);
        material < 0 restores the materials' own coefficients."""
        e = None if eval6 is None else np.ascontiguousarray(eval6, np.float64)
        f = None if frozen6 is None else np.ascontiguousarray(frozen6, np.float64)
        self._sss_keep = (e, f)
        self.lib.orc_diff_set_sss(self.handle, material, _p(e), _p(f))

    def diff_eval(self, views, shift=0, dimg=None, images=True, threads=8):
        """(loss, images): per-view mean images over frames frame0 + shift + s and
        the linear loss sum(dimg * sum_s L_s / spp)."""
        W, H, spp = views[0]["width"], views[0]["height"], views[0]["spp"]
        cams = np.array([list(v["eye"]) + list(v["U"]) + list(v["V"]) + list(v["W"]) for v in views], np.float32)
        f0 = np.array([v["frame0"] for v in views], np.uint32)
        img = np.zeros((len(views), H, W, 3)) if images else None
        dd = None if dimg is None else np.ascontiguousarray(dimg, np.float64)
        loss = self.lib.orc_diff_eval(self.handle, W, H, len(views), _p(cams), _p(f0), spp, shift, _p(dd), _p(img),
                                      threads)
        return loss, img

    # ------------------------------------------------ translucent material
    def probes_set(self, probes, samples, material):
        """Use the product's probe set (dict from rt.Context.probes()) with `samples`
        gather samples per hit on material handle `material`."""
        self._probe_keep = [np.ascontiguousarray(probes["pos"], np.float32),
                            np.ascontiguousarray(probes["nrm"], np.float32)]
        pos, nrm = self._probe_keep
        self.lib.orc_probes_set(self.handle, len(pos), _p(pos), _p(nrm), probes["cell"], samples, material)

    def probe_pass(self, frame, threads=8, normals=False):
        """(L, w12) of the probe pass for sample index `frame` (+ the entry normals sn)."""
        n = len(self._probe_keep[0])
        L = np.zeros((n, 4), np.float32)
        w = np.zeros((n, 4), np.float32)
        sn = np.zeros((n, 4), np.float32)
        self.lib.orc_probe_pass(self.handle, frame, _p(L), _p(w), _p(sn), threads)
        return (L, w, sn) if normals else (L, w)

    def probes_values(self, L, w12, sn=None):
        self._probe_vals = [np.ascontiguousarray(L, np.float32), np.ascontiguousarray(w12, np.float32),
                            None if sn is None else np.ascontiguousarray(sn, np.float32)]
        self.lib.orc_probes_values(self.handle, *[_p(a) for a in self._probe_vals])

    def sss_coefficients(self, material):
        out = np.zeros(64, np.float32)
        self.lib.orc_sss_coefficients(self.handle, material, _p(out))
        return out

    def dipole(self, material, r):
        r = np.ascontiguousarray(r, np.float32)
        out = np.zeros((len(r), 3), np.float32)
        self.lib.orc_dipole_eval(self.handle, material, len(r), _p(r), _p(out))
        return out

    def dirpole(self, material, rows):
        rows = np.ascontiguousarray(rows, np.float32)
        out = np.zeros((len(rows), 3), np.float32)
        self.lib.orc_dirpole_eval(self.handle, material, len(rows), _p(rows), _p(out))
        return out

    def dirpole_S_infinite(self, material, rows):
        rows = np.ascontiguousarray(rows, np.float64)
        out = np.zeros((len(rows), 3), np.float64)
        self.lib.orc_dirpole_S_infinite_eval(self.handle, material, len(rows), _p(rows), _p(out))
        return out

    def sss_gather(self, material, xo, no, reps, seed=1):
        xo = np.ascontiguousarray(xo, np.float32)
        no = np.ascontiguousarray(no, np.float32)
        out = np.zeros(3)
        self.lib.orc_sss_gather_eval(self.handle, material, _p(xo), _p(no), reps, seed, _p(out))
        return out
