// orf_diff_host.h -- host entry of the differentiable kernels (internal).
#pragma once

#include "orf_diff.h"
#include "orf_wavefront.h"

namespace orf {

// Samples of one pixel per work item of the differentiable kernels.
constexpr unsigned kDiffChunk = 2;

// Partial-sum entries a primal launch needs (float4 each).
inline long long diff_partial_entries(int nviews, int width, int height, unsigned spp)
{
    return (long long)nviews * width * height * (long long)((spp + kDiffChunk - 1) / kDiffChunk);
}

// One persistent launch over `nviews` views of width x height with `spp`
// samples each: primal (images, through `partial`: diff_partial_entries float4)
// or adjoint (dimages -> g_albedo, g_rough, g_sss).  counters: 8 device
// statistics zeroed at each launch, [0] dropped non-finite samples, [1] (adjoint)
// samples whose replay ended on a radiance other than their pass 1's; qctr: 8
// device work-queue counters.
void launch_diff(const SceneGPU& sc, const DiffParams& dp, const WfRender* d_views, int nviews, int width, int height,
                 unsigned spp, bool adjoint, unsigned frame_shift, float* images, float4* partial, const float* dimages,
                 float* g_albedo, float* g_rough, float* g_sss, unsigned long long* counters,
                 unsigned long long* qctr, hipStream_t stream);
// Recompute entry m of a device ScatteringGPU table from device parameters
// (absorption RGB, scattering RGB in 1/mm) times `scale`.
void launch_sss_coeffs(const float* params, float scale, ScatteringGPU* table, int m, hipStream_t stream);

}  // namespace orf
