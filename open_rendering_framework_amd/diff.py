"""Differentiable renders: a torch autograd Function (libtorch C++ extension)
over the C-ABI's path-replay adjoint.

build_extension() compiles the extension in-tree with one g++ command;
load_extension() imports the built module and fails loudly if it is missing.
"""
import glob
import importlib.util
import os

from . import INCLUDE_DIR, PACKAGE_DIR

EXT_NAME = "orf_torch_autograd"
BUILD_DIR = os.path.join(PACKAGE_DIR, "_torch_build")
SOURCE = os.path.join(PACKAGE_DIR, "torch_ext", "orf_autograd.cpp")

_ext = None


def build_extension(verbose=False):
    """Compile torch_ext/orf_autograd.cpp into BUILD_DIR with one explicit g++
    command (host C++ against libtorch and the HIP runtime; no source rewriting)."""
    import subprocess
    import sysconfig

    import torch
    import torch.utils.cpp_extension as ce

    os.makedirs(BUILD_DIR, exist_ok=True)
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(BUILD_DIR, EXT_NAME + suffix)
    abi = int(bool(torch._C._GLIBCXX_USE_CXX11_ABI))
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
           "-DTORCH_API_INCLUDE_EXTENSION_H", "-DTORCH_EXTENSION_NAME=" + EXT_NAME,
           "-D_GLIBCXX_USE_CXX11_ABI=%d" % abi, "-I" + INCLUDE_DIR, "-I" + sysconfig.get_paths()["include"]]
    cmd += ["-I" + d for d in ce.include_paths("cuda")]
    cmd += [SOURCE, "-o", out + ".tmp"]
    cmd += ["-L" + d for d in ce.library_paths("cuda")]
    # liborf_rt.so is found next to the package, wherever the package sits (the
    # extension lives in BUILD_DIR, one level below it)
    rel = os.path.relpath(PACKAGE_DIR, BUILD_DIR)
    cmd += ["-L" + PACKAGE_DIR, "-lorf_rt", "-Wl,-rpath,$ORIGIN/" + rel,
            "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python", "-lamdhip64"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)


def load_extension():
    global _ext
    if _ext is not None:
        return _ext
    import torch  # noqa: F401  (libtorch must be loaded first)

    from . import rt  # noqa: F401  (liborf_rt.so bound to torch's HIP runtime)

    hits = glob.glob(os.path.join(BUILD_DIR, EXT_NAME + "*.so"))
    if not hits:
        raise ImportError("%s not built in %s; run open_rendering_framework_amd.diff.build_extension()"
                          % (EXT_NAME, BUILD_DIR))
    spec = importlib.util.spec_from_file_location(EXT_NAME, hits[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    _ext = mod
    return mod


class DiffRenderer:
    """Renders views of a context's scene as a differentiable function of an
    albedo texture (on a Lambert material), a roughness texture (on a GGX
    conductor) and, with `scattering_material`, the absorption and scattering
    coefficients of a translucent material (6 values, 1/mm):
    images = renderer(albedo[h,w,3], roughness[h,w] [, scattering[6]])."""

    def __init__(self, ctx, views, albedo_material, roughness_material, tex_size=(64, 64),
                 albedo_uv=(0, (-10.0, -10.0, 10.0, 10.0)), roughness_uv=(1, (0.0, 0.0, 0.0, 0.0)),
                 frame_shift=0x40000000, scattering_material=-1):
        ext = load_extension()
        self.ctx = ctx
        self.scattering_material = scattering_material
        desc = [albedo_material, albedo_uv[0], *albedo_uv[1], roughness_material, roughness_uv[0], *roughness_uv[1],
                tex_size[0], tex_size[1], scattering_material]
        vv = [[v["width"], v["height"], v["spp"], v["frame0"], *v["eye"], *v["U"], *v["V"], *v["W"]] for v in views]
        self.views = views
        self.tex_size = tex_size
        self.setup = ext.make_setup(ctx.handle.value, [float(x) for x in desc], [[float(x) for x in v] for v in vv],
                                    int(frame_shift))
        self._ext = ext

    def __call__(self, albedo, roughness, scattering=None):
        if self.scattering_material >= 0:
            assert scattering is not None, "the scattering coefficients are a parameter of this renderer"
            return self._ext.render_sss(albedo, roughness, scattering, self.setup)
        return self._ext.render(albedo, roughness, self.setup)

    def close(self):
        if self.setup:
            self._ext.free_setup(self.setup)
            self.setup = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
