"""Scene descriptions for the render workloads (host side, numpy).

A SceneDesc is the flattened form of what the reference's JSON loader builds
(ORF/OptixSceneLoader.cpp:105-177): meshes with one material each, the light
list, the environment map, the pinhole camera and the integrator settings.
`to_context` pushes it through the C-ABI into a render context.

The benchmark scene ("plain path tracer", 512x512, 64 spp, max depth 8):
  ground   : 2-triangle quad, x,z in [-10, 10] at y = -1, Lambert albedo 0.5
  sphere   : icosphere level 7 (327,680 triangles), radius 1 at the origin,
             displaced radially by 3D value noise (lattice seed 7, amplitude
             0.05, frequency 4); GGX rough conductor, alpha 0.3, ior 0.27+2.78i
  light    : 1x1 quad at y = 3 facing -y, radiance 10 (emitter + triangle light)
  env      : procedural 1024x512 FP32 sky gradient + Gaussian sun lobe
             (sigma 0.5 deg, peak 5e4, direction SUN_DIR), importance sampled
  camera   : eye (0, 1.5, 4) looking at the origin, up +y, vfov 40
"""
import math
from dataclasses import dataclass, field

import numpy as np

# sun at 35 deg elevation, azimuth 150 deg in the env convention
# dir = (sin th sin ph, cos th, sin th cos ph), th = 55 deg, ph = 150 deg
_TH, _PH = math.radians(55.0), math.radians(150.0)
SUN_DIR = (math.sin(_TH) * math.sin(_PH), math.cos(_TH), math.sin(_TH) * math.cos(_PH))


@dataclass
class Material:
    program: str                      # reference closest-hit program file, e.g. "metallic_shader.cu"
    floats: dict = field(default_factory=dict)   # name -> tuple of floats
    uints: dict = field(default_factory=dict)    # name -> tuple of ints
    ior_complex: tuple = None         # conductor ior: ((re, im), (re, im), (re, im))


@dataclass
class Mesh:
    verts: np.ndarray                 # (n, 3) float32
    tris: np.ndarray                  # (m, 3) int32
    material: int
    normals: np.ndarray = None        # (k, 3) float32
    nindex: np.ndarray = None         # (m, 3) int32 (None: = tris when normals are per vertex)


@dataclass
class Light:
    kind: str                         # "point" | "directional" | "triangles" | "disk" | "sphere"
    radiance: tuple
    position: tuple = None            # point light position; disk / sphere centre
    direction: tuple = None
    tri_verts: np.ndarray = None      # (t, 9) float32
    tri_normals: np.ndarray = None    # (t, 9) float32
    radius: float = None              # disk / sphere
    theta: float = 0.0                # disk normal (sin t sin p, cos t, sin t cos p), degrees
    phi: float = 0.0


@dataclass
class SceneDesc:
    meshes: list
    materials: list
    lights: list
    env: np.ndarray = None            # (h, w, 4) float32
    env_importance: bool = True
    eye: tuple = (0.0, 0.0, 0.0)
    U: tuple = (1.0, 0.0, 0.0)
    V: tuple = (0.0, 1.0, 0.0)
    W: tuple = (0.0, 0.0, -1.0)
    width: int = 512
    height: int = 512
    max_depth: int = 8
    scene_epsilon: float = 1e-4
    ctx_uints: dict = field(default_factory=dict)  # context variables, e.g. probe_count

    def flat(self):
        """World-space triangle soup: verts, tris, normals, nindex, tri_mat."""
        V, T, N, NI, M = [], [], [], [], []
        vb = nb = 0
        for m in self.meshes:
            V.append(m.verts.astype(np.float32))
            T.append(m.tris.astype(np.int32) + vb)
            M.append(np.full(len(m.tris), m.material, np.int32))
            if m.normals is not None:
                N.append(m.normals.astype(np.float32))
                ni = m.nindex if m.nindex is not None else m.tris
                NI.append(ni.astype(np.int32) + nb)
                nb += len(m.normals)
            else:
                NI.append(np.full((len(m.tris), 3), -1, np.int32))
            vb += len(m.verts)
        N = np.concatenate(N) if N else np.zeros((0, 3), np.float32)
        return (np.concatenate(V), np.concatenate(T), N, np.concatenate(NI), np.concatenate(M))


def pinhole(eye, lookat, up, vfov, width, height):
    """U, V, W of PinholeCamera::calculateCameraVariables (ORF/Camera.cpp:83-103)."""
    eye = np.asarray(eye, np.float64)
    W = np.asarray(lookat, np.float64) - eye
    wlen = np.linalg.norm(W)
    U = np.cross(W, up)
    U /= np.linalg.norm(U)
    V = np.cross(U, W)
    V /= np.linalg.norm(V)
    vlen = wlen * math.tan(0.5 * vfov * math.pi / 180.0)
    V *= vlen
    U *= vlen * (width / height)
    f = lambda x: tuple(float(np.float32(c)) for c in x)  # noqa: E731
    return f(eye), f(U), f(V), f(W)


def quad(y, half, facing_down):
    """Axis-aligned square at height y; returns verts, tris with the given facing."""
    v = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], np.float32)
    if facing_down:
        t = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    else:
        t = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    return v, t


def icosphere(level=7, seed=7, amplitude=0.05, frequency=4.0):
    from . import rt

    return rt.gen_icosphere(level, seed, amplitude, frequency)


def sky_envmap(width=1024, height=512, sun_dir=SUN_DIR, sigma_deg=0.5, peak=5e4):
    from . import rt

    return rt.gen_sky_envmap(width, height, sun_dir, sigma_deg, peak)


def path_scene(width=512, height=512, level=7, ground_mat="lambert", sphere_mat="conductor", env=True,
               env_size=(1024, 512), max_depth=8, alpha=0.3, env_importance=True):
    """The plain-path-tracer benchmark scene (module docstring); smaller `level`,
    image and env sizes give the same scene family at test scale."""
    mats = [
        Material("diffuse_shader.cu", floats={"diffuse_color": (0.5, 0.5, 0.5)}),
        Material("metallic_shader.cu", floats={"roughness": (alpha, alpha)},
                 uints={"microfacet_model": (1,), "normal_distribution": (1,)},
                 ior_complex=((0.27, 2.78),) * 3),
        Material("arealight_shader.cu", floats={"radiance": (10.0, 10.0, 10.0)}),
    ]
    prog = {"lambert": 0, "conductor": 1}
    gv, gt = quad(-1.0, 10.0, facing_down=False)
    ground = Mesh(gv, gt, prog[ground_mat], normals=np.tile([[0, 1, 0]], (4, 1)).astype(np.float32))
    sv, sn, st = icosphere(level)
    sphere = Mesh(sv, st, prog[sphere_mat], normals=sn)
    lv, lt = quad(3.0, 0.5, facing_down=True)
    ln = np.tile([[0, -1, 0]], (4, 1)).astype(np.float32)
    light_mesh = Mesh(lv, lt, 2, normals=ln)
    tri_v = lv[lt].reshape(-1, 9)
    tri_n = ln[lt].reshape(-1, 9)
    lights = [Light("triangles", (10.0, 10.0, 10.0), tri_verts=tri_v, tri_normals=tri_n)]
    eye, U, V, W = pinhole((0.0, 1.5, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, width, height)
    envmap = sky_envmap(*env_size) if env else None
    return SceneDesc([ground, sphere, light_mesh], mats, lights, envmap, env_importance, eye, U, V, W,
                     width, height, max_depth, 1e-4)


class Pcg32:
    """PCG32 (O'Neill 2014, pcg-c-basic): initstate/initseq seeding, XSH-RR output."""

    MASK = (1 << 64) - 1

    def __init__(self, initstate, initseq):
        self.state = 0
        self.inc = ((initseq << 1) | 1) & self.MASK
        self.next()
        self.state = (self.state + initstate) & self.MASK
        self.next()

    def next(self):
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & self.MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self):
        return (self.next() >> 8) * (1.0 / 16777216.0)


BATCH_SPP = (4, 16, 64, 256)
BATCH_MATERIALS = ("lambert", "conductor", "dielectric")


def _round8(x):
    return int(min(1024, max(64, 8 * round(x / 8.0))))


def batch_workload(n=64, seed=1234, level=7, env_size=(1024, 512), max_depth=8, spp_scale=1.0, size_scale=1.0):
    """The batched workload: `n` renders of the path-tracer scene family.

    Render r draws from PCG32(seed, r): frame0 (its seed), width and height
    2^U[6,10] rounded to a multiple of 8, spp uniform in {4, 16, 64, 256}, and
    the sphere's material uniform among Lambert (albedo 0.5), GGX conductor
    (alpha 0.3, ior 0.27+2.78i) and Beckmann rough dielectric (alpha 0.3,
    ior 1.5, no absorption).  spp_scale / size_scale shrink the batch for tests.
    Returns (SceneDesc, items, remap) for Context.render_batch.
    """
    desc = path_scene(width=512, height=512, level=level, env_size=env_size, max_depth=max_depth)
    desc.materials.append(Material("rough_transparent_shader.cu",
                                   floats={"roughness": (0.3,), "ior": (1.5,), "glass_absorption": (0.0, 0.0, 0.0)},
                                   uints={"microfacet_model": (1,), "normal_distribution": (0,)}))
    desc.materials.append(Material("diffuse_shader.cu", floats={"diffuse_color": (0.5, 0.5, 0.5)}))
    sphere_handle = {"lambert": 4, "conductor": 1, "dielectric": 3}
    items, remap = [], []
    for r in range(n):
        g = Pcg32(seed, r)
        frame0 = g.next()
        w = _round8(2.0 ** (6.0 + 4.0 * g.uniform()) * size_scale)
        h = _round8(2.0 ** (6.0 + 4.0 * g.uniform()) * size_scale)
        spp = BATCH_SPP[min(3, int(4.0 * g.uniform()))]
        mat = BATCH_MATERIALS[min(2, int(3.0 * g.uniform()))]
        spp = max(1, int(round(spp * spp_scale)))
        eye, U, V, W = pinhole((0.0, 1.5, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, w, h)
        items.append({"width": w, "height": h, "spp": spp, "frame0": frame0, "eye": eye, "U": U, "V": V, "W": W,
                      "material": mat})
        remap.append([0, sphere_handle[mat], 2, 3, 4])
    return desc, items, np.array(remap, np.int32)


def subsurface_scene(width=512, height=512, level=8, glass_level=6, env_size=(1024, 512), max_depth=8,
                     probes=1 << 20, probe_samples=4, dipole_model=0, rough_boundary=None):
    """The subsurface + multiple-scattering scene: the path-tracer scene family with
    a displaced icosphere (radius 0.9, at x = -1) of whole milk (dipole BSSRDF,
    sigma_a = (0.0011, 0.0024, 0.014)/mm, sigma_s' = (2.55, 3.21, 3.77)/mm, eta 1.3,
    scale 100: one scene unit is 100 mm) next to a rough glass sphere (radius 0.7,
    at x = +1.1; GGX alpha 0.6, ior 1.5, microsurface multiple-scattering walk),
    ground, area light and sky.  `probes` Poisson-disk surface probes on the milk.
    rough_boundary = (microfacet_model, normal_distribution, roughness) gives the
    milk a microfacet boundary (rough_subsurface_scattering_shader.cu)."""
    milk_floats = {"absorption": (0.0011, 0.0024, 0.014), "scattering": (2.55, 3.21, 3.77),
                   "meancosine": (0.0, 0.0, 0.0), "scale": (100.0,), "ior": (1.3,)}
    milk_uints = {"dipole_model": (dipole_model,)}
    milk_program = "subsurface_scattering_shader.cu"
    if rough_boundary is not None:
        model, dist, alpha = rough_boundary
        milk_program = "rough_subsurface_scattering_shader.cu"
        milk_floats["roughness"] = (alpha, alpha)
        milk_uints.update({"microfacet_model": (model,), "normal_distribution": (dist,)})
    mats = [
        Material("diffuse_shader.cu", floats={"diffuse_color": (0.5, 0.5, 0.5)}),
        Material(milk_program, floats=milk_floats, uints=milk_uints),
        Material("arealight_shader.cu", floats={"radiance": (10.0, 10.0, 10.0)}),
        Material("rough_transparent_shader.cu",
                 floats={"roughness": (0.6,), "ior": (1.5,), "glass_absorption": (0.0, 0.0, 0.0)},
                 uints={"microfacet_model": (2,), "normal_distribution": (1,)}),
    ]
    gv, gt = quad(-1.0, 10.0, facing_down=False)
    ground = Mesh(gv, gt, 0, normals=np.tile([[0, 1, 0]], (4, 1)).astype(np.float32))
    sv, sn, st = icosphere(level)
    milk = Mesh((sv * 0.9 + np.array([-1.0, -0.1, 0.0], np.float32)).astype(np.float32), st, 1, normals=sn)
    gv2, gn2, gt2 = icosphere(glass_level)
    glass = Mesh((gv2 * 0.7 + np.array([1.1, -0.3, 0.0], np.float32)).astype(np.float32), gt2, 3, normals=gn2)
    lv, lt = quad(3.0, 0.5, facing_down=True)
    ln = np.tile([[0, -1, 0]], (4, 1)).astype(np.float32)
    light_mesh = Mesh(lv, lt, 2, normals=ln)
    lights = [Light("triangles", (10.0, 10.0, 10.0), tri_verts=lv[lt].reshape(-1, 9), tri_normals=ln[lt].reshape(-1, 9))]
    eye, U, V, W = pinhole((0.0, 1.5, 4.5), (0.0, -0.2, 0.0), (0.0, 1.0, 0.0), 45.0, width, height)
    envmap = sky_envmap(*env_size)
    return SceneDesc([ground, milk, glass, light_mesh], mats, lights, envmap, True, eye, U, V, W, width, height,
                     max_depth, 1e-4, ctx_uints={"probe_count": (probes,), "probe_samples": (probe_samples,)})


def ring_views(n=8, size=256, spp=16, radius=4.0, height=1.5, frame0=0, vfov=40.0):
    """`n` pinhole views on a ring of `radius` at `height`, looking at the origin;
    view i starts at frame frame0 + i * spp."""
    views = []
    for i in range(n):
        a = 2.0 * math.pi * i / n
        eye = (radius * math.sin(a), height, radius * math.cos(a))
        e, U, V, W = pinhole(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), vfov, size, size)
        views.append({"width": size, "height": size, "spp": spp, "frame0": frame0 + i * spp,
                      "eye": e, "U": U, "V": V, "W": W})
    return views


def diff_textures(tex=64, seed=99, albedo_range=(0.2, 0.8), rough_range=(0.1, 0.6)):
    """Random albedo (tex, tex, 3) and roughness (tex, tex) textures, PCG32(seed, 1/2)."""
    ga, gr = Pcg32(seed, 1), Pcg32(seed, 2)
    a = np.array([ga.uniform() for _ in range(tex * tex * 3)], np.float32).reshape(tex, tex, 3)
    r = np.array([gr.uniform() for _ in range(tex * tex)], np.float32).reshape(tex, tex)
    a = albedo_range[0] + (albedo_range[1] - albedo_range[0]) * a
    r = rough_range[0] + (rough_range[1] - rough_range[0]) * r
    return a.astype(np.float32), r.astype(np.float32)


# Differentiable form of path_scene: albedo texture on the Lambert ground
# (planar x,z map over the 20 x 20 quad), roughness texture on the conductor
# sphere (spherical map about the origin).
DIFF_ALBEDO_MATERIAL = 0
DIFF_ROUGHNESS_MATERIAL = 1
DIFF_ALBEDO_UV = (0, (-10.0, -10.0, 10.0, 10.0))
DIFF_ROUGHNESS_UV = (1, (0.0, 0.0, 0.0, 0.0))


def to_context(desc, ctx=None):
    """Create (or fill) a render context from a SceneDesc through the C-ABI."""
    from . import rt

    ctx = ctx or rt.Context()
    ctx.set_int("max_depth", desc.max_depth)
    for k, v in desc.ctx_uints.items():
        ctx.set_uint(k, *v)
    ctx.set_float("scene_epsilon", desc.scene_epsilon)
    for name in ("eye", "U", "V", "W"):
        ctx.set_float(name, *getattr(desc, name))
    handles = []
    for m in desc.materials:
        h = ctx.material(m.program)
        for k, v in m.floats.items():
            ctx.set_float(k, *v, material=h)
        for k, v in m.uints.items():
            ctx.set_uint(k, *v, material=h)
        if m.ior_complex is not None:
            ctx.set_conductor_ior(h, m.ior_complex)
        handles.append(h)
    for m in desc.meshes:
        ctx.mesh(m.verts, m.tris, handles[m.material], m.normals, m.nindex)
    for l in desc.lights:
        if l.kind == "point":
            ctx.point_light(l.position, l.radiance)
        elif l.kind == "directional":
            ctx.directional_light(l.direction, l.radiance)
        elif l.kind == "disk":
            ctx.disk_light(l.position, l.radius, l.theta, l.phi, l.radiance)
        elif l.kind == "sphere":
            ctx.sphere_light(l.position, l.radius, l.radiance)
        else:
            ctx.triangle_light(l.tri_verts, l.tri_normals, l.radiance)
    if desc.env is not None:
        ctx.envmap(desc.env, desc.env_importance)
    return ctx
