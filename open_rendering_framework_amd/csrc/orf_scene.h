// orf_scene.h -- device-side scene layout shared by the render kernels.
//
// HBM layout (all arrays 16 B aligned, structure-of-arrays where a kernel
// streams one field):
//   bvh4     : 4-wide BVH nodes, 128 B each: the four child boxes as six SoA
//              rows (lo.x, hi.x, lo.y, hi.y, lo.z, hi.z) and the four child links
//   tri      : triangles in leaf order, 48 B each: p0, e1 = p1-p0, e2 = p2-p0 and the
//              original triangle id (3 x float4), read by the intersection loop only
//   shade    : per original triangle: vertex-normal indices and material id,
//              read once per closest hit
//   normals  : per-vertex shading normals (float3 as float4)
//   materials, lights, triangle lights: small tables (read-only, cached)
//   env      : RGBA FP32 texture + FP32 marginal/conditional pdf and cdf tables
//   film     : FP64 RGBA per pixel: running sums of RGB and the finite-sample count
#pragma once

#include "orf_fresnel.h"
#include "orf_vec.h"

namespace orf {

// Material program ids (the reference binds closest-hit programs by file name).
enum MaterialKind : int {
    MAT_LAMBERT = 0,        // diffuse_shader.cu
    MAT_CONDUCTOR = 1,      // metallic_shader.cu
    MAT_EMITTER = 2,        // arealight_shader.cu
    MAT_DIELECTRIC = 3,     // rough_transparent_shader.cu
    MAT_INTERFACE = 4,      // lambertian_interface_shader.cu
    MAT_TRANSLUCENT = 5,    // subsurface_scattering_shader.cu (smooth boundary)
    MAT_ROUGH_TRANSLUCENT = 6,  // rough_subsurface_scattering_shader.cu (microfacet boundary)
    MAT_TRANSPARENT = 7,    // transparent_shader.cu (smooth glass)
    MAT_NORMAL = 8,         // normal_shader.cu (shading-normal display)
    MAT_FLAT = 9,           // flat_shader.cu (flat colour, optional shadow and highlight)
    MAT_ROUGH_DIFFUSE = 10, // rough_diffuse_shader.cu (Lambert base under a rough dielectric)
    MAT_COUNT
};

// Microfacet models and distributions (ORF/structs.h:242-266).
enum MicrofacetModel : int { WALTER_MODEL = 0, VISIBLE_NORMALS_MODEL = 1, MULTISCATTERING_MODEL = 2 };
enum NormalDistribution : int { BECKMANN_DISTRIBUTION = 0, GGX_DISTRIBUTION = 1 };

struct MaterialGPU {
    int kind;
    int microfacet_model;
    int normal_distribution;
    int sss;            // translucent: index into SceneGPU::sss (-1: none)
    f3 diffuse_color;   // Lambert albedo / interface albedo
    f3 radiance;        // emitter radiance / flat highlight colour
    f3 glass_absorption;
    f2 roughness;       // (ax, ay)
    float ior;          // dielectric ior (inside / outside)
    int flags;          // flat: kFlatShadow | kFlatHighlight
    ConductorEta eta;   // conductor: n1 / ior per channel
    float threshold;    // flat: highlight luminance threshold
};

constexpr int kFlatShadow = 1, kFlatHighlight = 2;

// Light types (ORF/structs.h:146-153) + the environment as a sampled light.
enum LightType : unsigned { POINT_LIGHT = 0, DIRECTIONAL_LIGHT = 1, TRIANGLES_AREA_LIGHT = 2,
                            DISK_LIGHT = 3, SPHERICAL_LIGHT = 4, ENV_LIGHT = 5 };

// LightStruct (ORF/structs.h:42-48): type, radiance, 3 parameters; the
// triangle-light form stores (count, start, end) as floats (:66-74).
struct LightGPU {
    unsigned type;
    f3 radiance;
    f3 params;
    f3 pad;
};

// TriangleLight (ORF/structs.h:95-102).
struct TriLightGPU {
    f3 v0, v1, v2;
    f3 n0, n1, n2;
    unsigned has_normals;
    f3 emission;
    float area;
};

struct EnvGPU {
    const float4* tex;   // w*h texels, row j <-> v = (j + 0.5) / h, theta = (1 - v) pi
    int w, h;
    const float* marg_pdf;  // h
    const float* marg_cdf;  // h
    const float* cond_pdf;  // h rows of w
    const float* cond_cdf;  // h rows of w
    int sampled;            // importance sampling tables valid
};

struct CameraGPU {
    f3 eye, U, V, W;
    unsigned patch_origin[2];
    unsigned patch_dims[2];
};

// ScatteringMaterialProperties (ORF/structs.h:116-144), FP32, derived on the host
// as ORF/RoughTranslucentMaterial.cpp:366-395 does (outside ior 1).
struct ScatteringGPU {
    float relative_ior;
    int dipole_model;   // 0 standard, 1 directional (ORF/structs.h:311-316)
    f3 absorption, scattering, meancosine;
    f3 extinction, reducedExtinction, deltaEddExtinction, D, transport;
    f3 C_phi, C_phi_inv, C_E, A, albedo, reducedAlbedo, de, three_D, two_A_de, global_coeff, one_over_three_ext;
    float mean_transport;
};

// Surface probes of the translucent object: Poisson-disk positions fixed at
// build time, per-launch irradiance from the probe pass, hashed uniform grid.
struct ProbesGPU {
    const float4* pos;   // xyz
    const float4* nrm;   // xyz (shading normal at the probe)
    float4* L;           // RGB: incident radiance estimate Le of this launch's probe sample
    float4* w12;         // xyz: refracted incident direction, w: T12 (times the rough boundary's weight)
    float4* sn;          // xyz: normal the light entered through (shading normal, or the
                         //      sampled microfacet normal of a rough boundary)
    const uint2* cells;  // hash slot -> (first probe, count); probes sorted by slot
    unsigned table_mask; // slots - 1 (power of two)
    float cell;          // grid cell size (>= the nearest-probe search radius)
    int count;           // probes (0: none)
    int samples;         // exit-radius samples per entering hit
    int sss;             // ScatteringGPU index of the probes' material (-1: none)
    int pad;
};

struct SceneGPU {
    const float4* bvh4;       // 8 float4 per node (the binned-SAH tree collapsed to 4 children; every traversal)
    int nodes4;               // 4-wide nodes (the first min(nodes4, kBvhTopOrdered) breadth-first)
    const float4* tri;        // 3 float4 per leaf-ordered triangle
    const int4* shade;        // per original triangle: nidx.xyz, material id
    const float4* normals;    // per normal
    const MaterialGPU* materials;
    const LightGPU* lights;
    int nlights;              // including the environment entry when sampled
    int root_is_leaf;         // tiny scenes: the root is a single leaf
    const TriLightGPU* tri_lights;
    // Analytic emitters of the disk and sphere lights (CU/disk_mesh.cu,
    // CU/sphere_mesh.cu): 2 float4 each, (centre, radius), (disk normal, kind
    // 0 disk / 1 sphere); tested after the BVH by every traversal.  A hit on
    // emitter i reports prim = analytic_prim + i (its shade record) and leaf = -1 - i.
    const float4* analytic;
    int nanalytic;
    int analytic_prim;
    EnvGPU env;
    int has_env;
    int max_depth;
    float scene_epsilon;
    unsigned features;        // OR of material_features() over the material table
    const ScatteringGPU* sss;  // per translucent material
    ProbesGPU probes;
};

// Material code a megakernel variant is compiled with: Lambert, conductor and
// emitter always; kFeatDielectric adds the rough dielectric (with the
// microsurface walk), the Lambertian interface, smooth glass, the normal
// display, the flat and the rough diffuse materials, kFeatTranslucent the
// translucent material.  Launches pick the smallest variant covering the scene,
// so simple scenes do not pay the register pressure of unused materials.
constexpr unsigned kFeatDielectric = 1u, kFeatTranslucent = 2u, kFeatAll = 3u;

inline bool is_translucent(int kind) { return kind == MAT_TRANSLUCENT || kind == MAT_ROUGH_TRANSLUCENT; }

inline unsigned material_features(int kind)
{
    return kind == MAT_DIELECTRIC || kind == MAT_INTERFACE || kind == MAT_TRANSPARENT || kind == MAT_NORMAL ||
                   kind == MAT_FLAT || kind == MAT_ROUGH_DIFFUSE
               ? kFeatDielectric
               : (kind == MAT_TRANSLUCENT || kind == MAT_ROUGH_TRANSLUCENT ? kFeatTranslucent : 0u);
}

constexpr float kRayMax = 1e27f;  // RT_DEFAULT_MAX
// The BVH builder lays out the first kBvhTopOrdered 4-wide nodes breadth-first
// (the top levels, larger boxes first within a level), so "node < n" selects the
// n nodes rays visit most; kernels copy such a prefix into LDS (kTopFloat4 float4
// per node: the six box rows and the links).
constexpr int kBvhTopOrdered = 1024;
constexpr int kTopFloat4 = 7;
constexpr int kStackDepth = 48;   // builder caps BVH depth below this
// Traversal stack entries per lane kept in LDS; deeper entries spill to a
// private (scratch) array.  Small LDS columns keep traversal kernels at high
// occupancy (LDS per wave64 = 256 B * kLdsStack).  On the 327,684-triangle
// displaced-icosphere scene 99.3% of camera, bounce and shadow rays of the
// 4-wide BVH never hold more than 8 entries (13 at most).
#ifndef ORF_LDS_STACK
#define ORF_LDS_STACK 8
#endif
constexpr int kLdsStack = ORF_LDS_STACK;
static_assert(kLdsStack >= 1 && kLdsStack <= kStackDepth && (kLdsStack & (kLdsStack - 1)) == 0,
              "LDS stack size must be a power of two no larger than kStackDepth");

// Triangle record word 3 of the first float4: the original triangle id, with
// kTriNoShadow set for triangles of a material whose shadow any-hit does not
// block (normal_shader.cu:18-22 terminates without clearing the attenuation).
constexpr unsigned kTriNoShadow = 0x40000000u, kTriIdMask = 0x3FFFFFFFu;

// BVH child link: internal node index, or leaf = 0x80000000 | (count-1) << 27 | first
ORF_HD bool link_is_leaf(unsigned c) { return (c & 0x80000000u) != 0; }
ORF_HD unsigned leaf_first(unsigned c) { return c & 0x07FFFFFFu; }
ORF_HD unsigned leaf_count(unsigned c) { return ((c >> 27) & 0xFu) + 1u; }

}  // namespace orf
