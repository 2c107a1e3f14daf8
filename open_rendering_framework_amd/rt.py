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
    "orf_context_create": [ctypes.c_int, ctypes.POINTER(c_vp)],
    "orf_context_destroy": [c_vp],
    "orf_variable_set_float": [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, c_vp, ctypes.c_int],
    "orf_variable_set_int": [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, c_vp, ctypes.c_int],
    "orf_variable_set_uint": [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, c_vp, ctypes.c_int],
    "orf_variable_set_user_data": [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, c_vp, ctypes.c_size_t],
    "orf_material_create": [c_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)],
    "orf_mesh_create": [c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "orf_mesh_set_transform": [c_vp, ctypes.c_int, c_vp],
    "orf_light_add_point": [c_vp, c_vp, c_vp],
    "orf_light_add_directional": [c_vp, c_vp, c_vp],
    "orf_light_add_triangles": [c_vp, c_vp, c_vp, c_i64, c_vp],
    "orf_light_add_disk": [c_vp, c_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, c_vp, c_vp],
    "orf_light_add_sphere": [c_vp, c_vp, ctypes.c_float, c_vp, c_vp],
    "orf_lights_clear": [c_vp],
    "orf_envmap_set": [c_vp, c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "orf_envmap_tables": [c_vp, c_vp, c_vp, c_vp, c_vp],
    "orf_accel_mark_dirty": [c_vp],
    "orf_accel_build": [c_vp, c_vp],
    "orf_launch": [c_vp, ctypes.c_int, ctypes.c_int, c_u32, c_u32, c_vp],
    "orf_launch_noalloc": [c_vp, ctypes.c_int, ctypes.c_int, c_u32, c_u32, c_vp, c_vp, c_vp],
    "orf_variable_set_matrix4x4": [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, c_vp],
    "orf_buffer_set": [c_vp, ctypes.c_char_p, c_vp, ctypes.c_size_t, ctypes.c_size_t],
    "orf_material_set_program": [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p],
    "orf_texture_create": [c_vp, c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.POINTER(ctypes.c_int)],
    "orf_variable_set_texture": [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int],
    "orf_film_clear": [c_vp],
    "orf_film_read": [c_vp, c_vp, c_vp],
    "orf_film_read_sums": [c_vp, c_vp],
    "orf_counters": [c_vp, c_vp],
    "orf_film_display": [c_vp, c_vp, c_vp],
    "orf_diff_counters": [c_vp, c_vp],
    "orf_gen_icosphere": [ctypes.c_int, c_u32, ctypes.c_float, ctypes.c_float, c_vp, c_vp,
                          ctypes.POINTER(c_i64), c_vp, ctypes.POINTER(c_i64)],
    "orf_gen_sky_envmap": [ctypes.c_int, ctypes.c_int, c_vp, ctypes.c_float, ctypes.c_float, c_vp],
    "orf_wavefront_default_config": [c_vp],
    "orf_render_batch": [c_vp, c_vp, ctypes.c_int32, c_vp, ctypes.c_int32, ctypes.c_int32, c_vp, c_vp, c_vp],
    "orf_batch_film_read": [c_vp, ctypes.c_int32, c_vp],
    "orf_batch_film_sums": [c_vp, ctypes.c_int32, c_vp],
    "orf_batch_counters": [c_vp, c_vp],
    "orf_wavefront_autotune": [c_vp, c_vp, ctypes.c_int32, c_vp, ctypes.c_int32, c_u32, c_vp, ctypes.c_char_p,
                               ctypes.c_size_t],
    "orf_diff_render": [c_vp, c_vp, c_vp, ctypes.c_int32, c_vp, c_vp, c_vp, c_vp],
    "orf_scattering_properties": [c_vp, ctypes.c_int, c_vp, ctypes.c_int],
    "orf_probes_read": [c_vp, ctypes.POINTER(c_i64), c_dblp, c_f32p, c_f32p, c_vp, c_vp, c_vp, c_vp, c_vp],
    "orf_probe_pass": [c_vp, c_u32, c_vp],
    "orf_diff_adjoint": [c_vp, c_vp, c_vp, ctypes.c_int32, c_vp, c_vp, c_vp, c_vp, c_vp, c_u32, c_vp],
    "orf_diff_render_ex": [c_vp, c_vp, c_vp, ctypes.c_int32, c_vp, c_vp, c_vp, c_vp, c_vp],
    "orf_diff_adjoint_ex": [c_vp, c_vp, c_vp, ctypes.c_int32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u32, c_vp],
    "orf_scene_load_json": [c_vp, ctypes.c_char_p, c_u32, c_vp],
    "orf_integrator_set": [c_vp, ctypes.c_int],
    "orf_render_frames": [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, c_vp],
    "orf_film_save_raw": [c_vp, ctypes.c_char_p],
    "orf_film_load_raw": [c_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32],
    "orf_film_save_png": [c_vp, ctypes.c_char_p],
    "orf_scene_save": [c_vp, ctypes.c_char_p, ctypes.c_uint32],
    "orf_hdr_read": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), c_vp],
    "orf_obj_read": [ctypes.c_char_p, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                     ctypes.POINTER(ctypes.c_int32), c_vp, c_vp, c_vp, c_vp],
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


def _np_ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def gen_icosphere(level, seed, amplitude, frequency):
    """(verts (n,3) f32, normals (n,3) f32, tris (m,3) i32) of the displaced icosphere."""
    import numpy as np

    nv, nt = c_i64(0), c_i64(0)
    call("orf_gen_icosphere", level, seed, amplitude, frequency, None, None, ctypes.byref(nv), None, ctypes.byref(nt))
    v = np.zeros((nv.value, 3), np.float32)
    n = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.int32)
    call("orf_gen_icosphere", level, seed, amplitude, frequency, _np_ptr(v), _np_ptr(n), ctypes.byref(nv),
         _np_ptr(t), ctypes.byref(nt))
    return v, n, t


def gen_sky_envmap(width, height, sun_dir, sigma_deg, peak):
    import numpy as np

    out = np.zeros((height, width, 4), np.float32)
    sd = (ctypes.c_float * 3)(*sun_dir)
    call("orf_gen_sky_envmap", width, height, sd, sigma_deg, peak, _np_ptr(out))
    return out


def hdr_read(path):
    """Radiance RGBE image as (h, w, 4) FP32, row 0 = the file's bottom row."""
    import numpy as np

    w, h = ctypes.c_int32(), ctypes.c_int32()
    p = str(path).encode()
    call("orf_hdr_read", p, ctypes.byref(w), ctypes.byref(h), None)
    out = np.zeros((h.value, w.value, 4), np.float32)
    call("orf_hdr_read", p, ctypes.byref(w), ctypes.byref(h), _np_ptr(out))
    return out


def obj_read(path):
    """Wavefront OBJ: dict verts (n,3), tris (m,3), normals (k,3), nindex (m,3) or None."""
    import numpy as np

    nv, nt, nn, hn = c_i64(), c_i64(), c_i64(), ctypes.c_int32()
    p = str(path).encode()
    call("orf_obj_read", p, ctypes.byref(nv), ctypes.byref(nt), ctypes.byref(nn), ctypes.byref(hn),
         None, None, None, None)
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.int32)
    n = np.zeros((nn.value, 3), np.float32)
    ni = np.zeros((nt.value, 3), np.int32) if hn.value else None
    call("orf_obj_read", p, ctypes.byref(nv), ctypes.byref(nt), ctypes.byref(nn), ctypes.byref(hn),
         _np_ptr(v), _np_ptr(t), _np_ptr(n), _np_ptr(ni))
    return {"verts": v, "tris": t, "normals": n, "nindex": ni}


class SceneInfo(ctypes.Structure):
    """orf_scene_info."""
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("frame_count", ctypes.c_int32),
                ("meshes", ctypes.c_int32), ("materials", ctypes.c_int32), ("lights", ctypes.c_int32),
                ("buffer_path", ctypes.c_char * 512)]


LOAD_ENV_IMPORTANCE = 1
INTEGRATOR_PATH_TRACER, INTEGRATOR_DEPTH = 0, 1
SCOPE_CONTEXT, SCOPE_MATERIAL, SCOPE_GEOMETRY, SCOPE_INSTANCE = 0, 1, 2, 3
RAY_RADIANCE, RAY_SHADOW, RAY_DEPTH = 0, 1, 2
WRAP_REPEAT, WRAP_CLAMP_TO_EDGE = 0, 1
FILTER_LINEAR, FILTER_NEAREST = 0, 1


def _structs():
    import numpy as np

    f3 = ("<f4", (3,))
    light = np.dtype([("light_type", "<u4"), ("emitted_radiance", f3), ("light_parameters", f3), ("padding", f3)])
    tri = np.dtype([("v0", f3), ("v1", f3), ("v2", f3), ("n0", f3), ("n1", f3), ("n2", f3), ("has_normals", "<u4"),
                    ("emission", f3), ("area", "<f4")])
    names = ["relative_ior", "ior_real", "ior_imag", "absorption", "scattering", "meancosine", "extinction",
             "reducedExtinction", "deltaEddExtinction", "D", "transport", "C_phi", "C_phi_inv", "C_E", "A", "albedo",
             "reducedAlbedo", "de", "three_D", "two_A_de", "global_coeff", "one_over_three_ext", "mean_transport"]
    sss = np.dtype([(n, "<f4") if n in ("relative_ior", "mean_transport") else (n, f3) for n in names])
    assert light.itemsize == 40 and tri.itemsize == 92 and sss.itemsize == 260
    return light, tri, sss


# numpy layouts of LightStruct, TriangleLight and ScatteringMaterialProperties (ORF/structs.h)
LIGHT_STRUCT, TRIANGLE_LIGHT, SCATTERING_PROPERTIES = _structs()
BATCH_WAVEFRONT, BATCH_MEGAKERNEL, BATCH_PERSISTENT = 0, 1, 2


class RenderItem(ctypes.Structure):
    """orf_render_item: one render of a batch."""

    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("spp", ctypes.c_uint32),
                ("frame0", ctypes.c_uint32), ("eye", ctypes.c_float * 3), ("U", ctypes.c_float * 3),
                ("V", ctypes.c_float * 3), ("W", ctypes.c_float * 3)]


class WavefrontConfig(ctypes.Structure):
    """orf_wavefront_config: the wavefront's tunable parameters."""

    _fields_ = [("block_size", ctypes.c_int32), ("pool_size", ctypes.c_int32), ("sort_policy", ctypes.c_int32),
                ("regen_threshold", ctypes.c_float), ("use_graph", ctypes.c_int32), ("graph_pairs", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

    @classmethod
    def default(cls):
        c = cls()
        call("orf_wavefront_default_config", ctypes.byref(c))
        return c


class DiffDesc(ctypes.Structure):
    """orf_diff_desc: which materials carry the albedo / roughness textures."""

    _fields_ = [("albedo_material", ctypes.c_int32), ("albedo_uv", ctypes.c_int32),
                ("albedo_uv_params", ctypes.c_float * 4), ("roughness_material", ctypes.c_int32),
                ("roughness_uv", ctypes.c_int32), ("roughness_uv_params", ctypes.c_float * 4),
                ("tex_w", ctypes.c_int32), ("tex_h", ctypes.c_int32), ("scattering_material", ctypes.c_int32)]

    @classmethod
    def make(cls, albedo_material, roughness_material, tex_size, albedo_uv=(0, (-10.0, -10.0, 10.0, 10.0)),
             roughness_uv=(1, (0.0, 0.0, 0.0, 0.0)), scattering_material=-1):
        d = cls()
        d.albedo_material, d.roughness_material = albedo_material, roughness_material
        d.albedo_uv, d.roughness_uv = albedo_uv[0], roughness_uv[0]
        d.albedo_uv_params[:] = [float(x) for x in albedo_uv[1]]
        d.roughness_uv_params[:] = [float(x) for x in roughness_uv[1]]
        d.tex_w, d.tex_h = tex_size
        d.scattering_material = scattering_material
        return d


class Context:
    """A render context (C-ABI orf_context) on one device."""

    def __init__(self, device=0):
        h = c_vp()
        call("orf_context_create", device, ctypes.byref(h))
        self.handle = h

    def close(self):
        if self.handle:
            call("orf_context_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- variables
    def _scope(self, material, instance=None):
        if instance is not None:
            return SCOPE_INSTANCE, instance
        return (SCOPE_CONTEXT, 0) if material is None else (SCOPE_MATERIAL, material)

    def set_float(self, name, *v, material=None, instance=None):
        sc, h = self._scope(material, instance)
        arr = (ctypes.c_float * len(v))(*v)
        call("orf_variable_set_float", self.handle, sc, h, name.encode(), arr, len(v))

    def set_int(self, name, *v, material=None, instance=None):
        sc, h = self._scope(material, instance)
        arr = (ctypes.c_int32 * len(v))(*v)
        call("orf_variable_set_int", self.handle, sc, h, name.encode(), arr, len(v))

    def set_uint(self, name, *v, material=None, instance=None):
        sc, h = self._scope(material, instance)
        arr = (ctypes.c_uint32 * len(v))(*v)
        call("orf_variable_set_uint", self.handle, sc, h, name.encode(), arr, len(v))

    def set_user_data(self, name, data, material=None, instance=None):
        """setUserData(size, ptr): `data` is bytes (or a numpy array viewed as bytes)."""
        sc, h = self._scope(material, instance)
        b = bytes(data)
        buf = ctypes.create_string_buffer(b, len(b))
        call("orf_variable_set_user_data", self.handle, sc, h, name.encode(), buf, len(b))

    def set_matrix4x4(self, name, m4, transpose=False, instance=None):
        """setMatrix4x4fv(transpose, m) on the context or a mesh instance."""
        import numpy as np

        m = np.ascontiguousarray(m4, np.float32).reshape(16)
        sc, h = self._scope(None, instance)
        call("orf_variable_set_matrix4x4", self.handle, sc, h, name.encode(), int(bool(transpose)), _np_ptr(m))

    def set_buffer(self, name, records):
        """A user-format context buffer from a numpy structured array
        (LIGHT_STRUCT for light_buffer, TRIANGLE_LIGHT for triangle_light_buffer)."""
        import numpy as np

        a = np.ascontiguousarray(records)
        call("orf_buffer_set", self.handle, name.encode(), _np_ptr(a) if len(a) else None, a.dtype.itemsize, len(a))

    def material_program(self, material, ray_type, program_file, entry):
        call("orf_material_set_program", self.handle, material, ray_type, program_file.encode(), entry.encode())

    def texture(self, rgba, wrap=0, filter=0):
        import numpy as np

        t = np.ascontiguousarray(rgba, np.float32)
        h = ctypes.c_int()
        call("orf_texture_create", self.handle, _np_ptr(t), t.shape[1], t.shape[0], wrap, filter, ctypes.byref(h))
        return h.value

    def set_texture(self, name, texture):
        call("orf_variable_set_texture", self.handle, SCOPE_CONTEXT, 0, name.encode(), texture)

    def set_conductor_ior(self, material, ior):
        """ior: three (re, im) pairs, passed as MyComplex3 user data."""
        flat = [float(x) for pair in ior for x in pair]
        arr = (ctypes.c_double * 6)(*flat)
        call("orf_variable_set_user_data", self.handle, SCOPE_MATERIAL, material, b"ior", arr, 48)

    # ---- scene objects
    def material(self, program):
        h = ctypes.c_int()
        call("orf_material_create", self.handle, program.encode(), ctypes.byref(h))
        return h.value

    def mesh(self, verts, tris, material, normals=None, nindex=None):
        import numpy as np

        v = np.ascontiguousarray(verts, np.float32)
        t = np.ascontiguousarray(tris, np.int32)
        n = None if normals is None else np.ascontiguousarray(normals, np.float32)
        ni = None if nindex is None else np.ascontiguousarray(nindex, np.int32)
        h = ctypes.c_int()
        call("orf_mesh_create", self.handle, _np_ptr(v), len(v), _np_ptr(t), len(t), _np_ptr(n),
             0 if n is None else len(n), _np_ptr(ni), material, ctypes.byref(h))
        return h.value

    def set_transform(self, mesh, m4):
        import numpy as np

        m = np.ascontiguousarray(m4, np.float32).reshape(16)
        call("orf_mesh_set_transform", self.handle, mesh, _np_ptr(m))

    def point_light(self, position, radiance):
        call("orf_light_add_point", self.handle, (ctypes.c_float * 3)(*position), (ctypes.c_float * 3)(*radiance))

    def directional_light(self, direction, radiance):
        call("orf_light_add_directional", self.handle, (ctypes.c_float * 3)(*direction),
             (ctypes.c_float * 3)(*radiance))

    def triangle_light(self, tri_verts, tri_normals, radiance):
        import numpy as np

        tv = np.ascontiguousarray(tri_verts, np.float32).reshape(-1, 9)
        tn = None if tri_normals is None else np.ascontiguousarray(tri_normals, np.float32).reshape(-1, 9)
        call("orf_light_add_triangles", self.handle, _np_ptr(tv), _np_ptr(tn), len(tv),
             (ctypes.c_float * 3)(*radiance))

    def disk_light(self, position, radius, theta, phi, radiance):
        """DiskAreaLight (normal from theta, phi in degrees); returns the emitter's material handle."""
        m = ctypes.c_int()
        call("orf_light_add_disk", self.handle, (ctypes.c_float * 3)(*position), radius, theta, phi,
             (ctypes.c_float * 3)(*radiance), ctypes.byref(m))
        return m.value

    def sphere_light(self, position, radius, radiance):
        """SphericalLight; returns the emitter's material handle."""
        m = ctypes.c_int()
        call("orf_light_add_sphere", self.handle, (ctypes.c_float * 3)(*position), radius,
             (ctypes.c_float * 3)(*radiance), ctypes.byref(m))
        return m.value

    def envmap(self, rgba, importance_sample=True):
        import numpy as np

        e = np.ascontiguousarray(rgba, np.float32)
        h, w = e.shape[:2]
        call("orf_envmap_set", self.handle, _np_ptr(e), w, h, 1 if importance_sample else 0)

    def env_tables(self, w, h):
        import numpy as np

        out = [np.zeros(h, np.float32), np.zeros(h, np.float32), np.zeros(h * w, np.float32),
               np.zeros(h * w, np.float32)]
        call("orf_envmap_tables", self.handle, *[_np_ptr(a) for a in out])
        return out

    def accel_build(self):
        import numpy as np

        st = np.zeros(6, np.int64)
        call("orf_accel_build", self.handle, _np_ptr(st))
        return dict(zip(("nodes", "leaves", "max_depth", "max_leaf", "tris", "degenerate"), st.tolist()))

    # ---- launches and film
    def launch(self, width, height, frame0=0, spp=1, stream=None):
        call("orf_launch", self.handle, width, height, frame0, spp, stream_ptr(stream))

    def launch_noalloc(self, width, height, frame0, spp, film, counters=None, stream=None):
        """orf_launch into a caller-owned device film (float64 tensor of
        height*width*4 sums) and optional device counters (uint64/int64 tensor
        of 4); the call allocates no device memory."""
        assert film.is_cuda and film.dtype == torch.float64 and film.numel() == 4 * width * height
        assert film.is_contiguous()
        if counters is not None:
            assert counters.is_cuda and counters.element_size() == 8 and counters.numel() == 4
        call("orf_launch_noalloc", self.handle, width, height, frame0, spp, ptr(film), ptr(counters),
             stream_ptr(stream))

    def film_clear(self):
        call("orf_film_clear", self.handle)

    def film(self, width, height):
        import numpy as np

        out = np.zeros((height, width, 4), np.float64)
        call("orf_film_read", self.handle, _np_ptr(out), None)
        return out

    def film_sums(self, width, height):
        import numpy as np

        out = np.zeros((height, width, 4), np.float64)
        call("orf_film_read_sums", self.handle, _np_ptr(out))
        return out

    def film_display(self, out, stream=None):
        """Film mean RGBA (as film()) into the device float32 tensor `out`
        (height, width, 4) on `stream`, without synchronising."""
        call("orf_film_display", self.handle, ptr(out), stream_ptr(stream))
        return out

    def counters(self):
        import numpy as np

        out = np.zeros(4, np.uint64)
        call("orf_counters", self.handle, _np_ptr(out))
        return dict(zip(("nonfinite", "ext_rays", "shadow_rays", "paths"), [int(x) for x in out]))

    # ---- batched renders
    @staticmethod
    def _items(items):
        arr = (RenderItem * len(items))()
        for i, it in enumerate(items):
            arr[i].width, arr[i].height = it["width"], it["height"]
            arr[i].spp, arr[i].frame0 = it["spp"], it["frame0"]
            for k in ("eye", "U", "V", "W"):
                getattr(arr[i], k)[:] = list(it[k])
        return arr

    def render_batch(self, items, remap=None, mode=BATCH_WAVEFRONT, config=None, stream=None):
        """Render every item; returns the wall time in seconds."""
        import numpy as np

        arr = self._items(items)
        rm = None if remap is None else np.ascontiguousarray(remap, np.int32)
        stride = 0 if rm is None else rm.shape[1]
        secs = ctypes.c_double()
        call("orf_render_batch", self.handle, arr, len(items), _np_ptr(rm), stride, mode,
             None if config is None else ctypes.byref(config), stream_ptr(stream), ctypes.byref(secs))
        return secs.value

    def batch_film(self, item, width, height, sums=False):
        import numpy as np

        out = np.zeros((height, width, 4), np.float64)
        call("orf_batch_film_sums" if sums else "orf_batch_film_read", self.handle, item, _np_ptr(out))
        return out

    def batch_counters(self):
        import numpy as np

        out = np.zeros(4, np.uint64)
        call("orf_batch_counters", self.handle, _np_ptr(out))
        return dict(zip(("nonfinite", "ext_rays", "shadow_rays", "paths"), [int(x) for x in out]))

    def diff_counters(self):
        """Statistics of the last differentiable launch: non-finite samples and
        adjoint samples whose replay did not reproduce their first pass."""
        import numpy as np

        out = np.zeros(2, np.uint64)
        call("orf_diff_counters", self.handle, _np_ptr(out))
        return dict(zip(("nonfinite", "replay_mismatch"), [int(x) for x in out]))

    def diff_render(self, desc, views, albedo, roughness, images=None, stream=None, scattering=None):
        """Per-view mean images (nviews, h, w, 3) of the textured scene (device tensors);
        `scattering` (6 floats, device) when desc.scattering_material is set."""
        arr = self._items(views)
        if images is None:
            dev = next(t for t in (albedo, roughness, scattering) if t is not None).device
            images = torch.empty((len(views), views[0]["height"], views[0]["width"], 3), dtype=torch.float32,
                                 device=dev)
        call("orf_diff_render_ex", self.handle, ctypes.byref(desc), arr, len(views), ptr(albedo), ptr(roughness),
             ptr(scattering), ptr(images), stream_ptr(stream))
        return images

    def diff_adjoint(self, desc, views, albedo, roughness, d_images, g_albedo=None, g_roughness=None,
                     frame_shift=0x40000000, stream=None):
        """Adds dLoss/dtexture into (g_albedo, g_roughness) (zeroed here when not given)."""
        arr = self._items(views)
        if g_albedo is None and albedo is not None:
            g_albedo = torch.zeros_like(albedo)
        if g_roughness is None and roughness is not None:
            g_roughness = torch.zeros_like(roughness)
        call("orf_diff_adjoint", self.handle, ctypes.byref(desc), arr, len(views), ptr(albedo), ptr(roughness),
             ptr(d_images), ptr(g_albedo), ptr(g_roughness), frame_shift, stream_ptr(stream))
        return g_albedo, g_roughness

    def diff_adjoint_sss(self, desc, views, albedo, roughness, scattering, d_images, frame_shift=0x40000000,
                         stream=None):
        """dLoss/d(textures, scattering coefficients) by orf_diff_adjoint_ex (fresh zeroed buffers)."""
        arr = self._items(views)
        g_albedo = torch.zeros_like(albedo) if albedo is not None else None
        g_roughness = torch.zeros_like(roughness) if roughness is not None else None
        g_sss = torch.zeros(6, dtype=torch.float32, device=scattering.device)
        call("orf_diff_adjoint_ex", self.handle, ctypes.byref(desc), arr, len(views), ptr(albedo), ptr(roughness),
             ptr(scattering), ptr(d_images), ptr(g_albedo), ptr(g_roughness), ptr(g_sss), frame_shift,
             stream_ptr(stream))
        return g_albedo, g_roughness, g_sss

    SSS_FIELDS = ("absorption", "scattering", "meancosine", "extinction", "reducedExtinction",
                  "deltaEddExtinction", "D", "transport", "C_phi", "C_phi_inv", "C_E", "A", "albedo",
                  "reducedAlbedo", "de", "three_D", "two_A_de", "global_coeff", "one_over_three_ext")

    def scattering_properties(self, material):
        """Derived coefficients of a translucent material as a dict of floats / 3-tuples."""
        import numpy as np

        out = np.zeros(64, np.float32)
        call("orf_scattering_properties", self.handle, material, _np_ptr(out), 64)
        d = {"relative_ior": float(out[0]), "dipole_model": int(out[1])}
        for i, k in enumerate(self.SSS_FIELDS):
            d[k] = tuple(float(x) for x in out[2 + 3 * i: 5 + 3 * i])
        d["mean_transport"] = float(out[2 + 3 * len(self.SSS_FIELDS)])
        return d

    def probes(self):
        """Probe set: dict with count, area, cell, rmin and (n, 4) arrays pos, nrm, L, w12, sn."""
        import numpy as np

        n, area = c_i64(0), ctypes.c_double()
        cell, rmin = ctypes.c_float(), ctypes.c_float()
        call("orf_probes_read", self.handle, ctypes.byref(n), ctypes.byref(area), ctypes.byref(cell),
             ctypes.byref(rmin), None, None, None, None, None)
        arrs = [np.zeros((n.value, 4), np.float32) for _ in range(5)]
        call("orf_probes_read", self.handle, None, None, None, None, *[_np_ptr(a) for a in arrs])
        return {"count": n.value, "area": area.value, "cell": cell.value, "rmin": rmin.value,
                "pos": arrs[0], "nrm": arrs[1], "L": arrs[2], "w12": arrs[3], "sn": arrs[4]}

    def load_scene(self, path, env_importance=False):
        """Populate the context from a JSON scene file (orf_scene_load_json); returns its info dict."""
        info = SceneInfo()
        call("orf_scene_load_json", self.handle, str(path).encode(), LOAD_ENV_IMPORTANCE if env_importance else 0,
             ctypes.byref(info))
        return {"width": info.width, "height": info.height, "frame_count": info.frame_count, "meshes": info.meshes,
                "materials": info.materials, "lights": info.lights, "buffer_path": info.buffer_path.decode()}

    def integrator(self, kind):
        """INTEGRATOR_PATH_TRACER or INTEGRATOR_DEPTH (orf_integrator_set)."""
        call("orf_integrator_set", self.handle, kind)

    def render_frames(self, width, height, frame_begin, frame_end, stream=None):
        """One launch per frame in [frame_begin, frame_end) (orf_render_frames)."""
        call("orf_render_frames", self.handle, width, height, frame_begin, frame_end, stream_ptr(stream))

    def save_raw(self, path):
        call("orf_film_save_raw", self.handle, str(path).encode())

    def load_raw(self, path, width, height, frames):
        call("orf_film_load_raw", self.handle, str(path).encode(), width, height, frames)

    def save_png(self, path):
        call("orf_film_save_png", self.handle, str(path).encode())

    def save_scene(self, path, frame_count):
        call("orf_scene_save", self.handle, str(path).encode(), frame_count)

    def probe_pass(self, frame, stream=None):
        call("orf_probe_pass", self.handle, frame, stream_ptr(stream))

    def autotune(self, items, remap=None, spp_divisor=16):
        """(best WavefrontConfig, report dict) of the wavefront coordinate search."""
        import json

        import numpy as np

        arr = self._items(items)
        rm = None if remap is None else np.ascontiguousarray(remap, np.int32)
        stride = 0 if rm is None else rm.shape[1]
        best = WavefrontConfig()
        buf = ctypes.create_string_buffer(1 << 16)
        call("orf_wavefront_autotune", self.handle, arr, len(items), _np_ptr(rm), stride, spp_divisor,
             ctypes.byref(best), buf, len(buf))
        return best, json.loads(buf.value.decode())
