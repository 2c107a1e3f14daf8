// orf_probes_host.h -- translucent-material host data: scattering parameters,
// probe generation and the probe pass launch (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "orf_scene.h"

namespace orf {

// Parameters of a translucent material as the reference's variables name them
// (absorption, scattering, meancosine per RGB in 1/mm, scale, ior, dipole_model).
struct SssParamsHost {
    f3 absorption = mk3(0.0011f, 0.0024f, 0.014f);  // Wholemilk (ORF/RoughTranslucentMaterial.cpp:258-263)
    f3 scattering = mk3(2.55f, 3.21f, 3.77f);
    f3 meancosine = mk3(0.0f);
    float scale = 100.0f;  // ORF/RoughTranslucentMaterial.cpp:13
    float ior = 1.3f;
    int dipole_model = 0;
    // coefficients given whole as ScatteringMaterialProperties user data: used
    // as they are (dipole_model still from the material variable)
    bool has_user = false;
    ScatteringGPU user;
};

ScatteringGPU scattering_coefficients(const SssParamsHost& h);

struct ProbeSetHost {
    std::vector<float4> pos, nrm;  // sorted by grid slot
    std::vector<uint2> cells;      // slot -> (first, count)
    unsigned table_mask = 0;
    float cell = 0.0f;             // grid cell / nearest-probe search radius
    float rmin = 0.0f;             // Poisson-disk minimum distance
    double area = 0.0;             // surface area of the object
};

// Poisson-disk probes on the triangles of material `material` (world space).
ProbeSetHost generate_probes(const float* V, const int32_t* I, const int4* shade, const float4* normals,
                             int64_t ntris, int material, int count, uint32_t seed);

// Probe pass: one incident-light sample per probe for sample index `frame`.
// `sc` must carry the probes (their positions/normals/outputs) with gathers
// disabled (probes.sss = -1), so no probe reads another probe of this pass.
void launch_probe_pass(const SceneGPU& sc, const ScatteringGPU* d_props, int sss, unsigned frame,
                       hipStream_t stream);

}  // namespace orf
