// orf_render.h -- host-side entry points of the render kernels (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "orf_scene.h"

namespace orf {

struct FilmGPU {
    double* sums;                  // 4 per pixel: R, G, B sums and finite count
    unsigned long long* counters;  // [0] non-finite, [1] extension rays, [2] shadow rays, [3] paths
    int width, height;
};

// Megakernel: every (pixel, sample) path of frames frame0 .. frame0+spp-1.
// remap: per-render material table (device, nullptr = identity).
void launch_megakernel(const SceneGPU& sc, const CameraGPU& cam, const FilmGPU& film, const int* remap,
                       unsigned frame0, unsigned spp, hipStream_t stream);

// FP64 construction of the environment's FP32 pdf/cdf tables (spec CU/env_cameras.cu).
// scratch: (h + 1) doubles.
void build_env_tables(const float4* tex, int w, int h, float* marg_pdf, float* marg_cdf, float* cond_pdf,
                      float* cond_cdf, double* scratch, hipStream_t stream);

// Conductor constants as the reference forms eta = 1 / ior (ORF/MyComplex.h:155-160).
void conductor_eta_host(const double ior_re[3], const double ior_im[3], ConductorEta& e);

}  // namespace orf
