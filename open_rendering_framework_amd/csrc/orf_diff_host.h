// orf_diff_host.h -- host entry of the differentiable kernels (internal).
#pragma once

#include "orf_diff.h"
#include "orf_wavefront.h"

namespace orf {

// One launch over `nviews` views of width x height: primal (images) or adjoint
// (dimages -> g_albedo, g_rough, g_sss).  counters[0] receives dropped non-finite samples.
void launch_diff(const SceneGPU& sc, const DiffParams& dp, const WfRender* d_views, int nviews, int width,
                 int height, bool adjoint, unsigned frame_shift, float* images, const float* dimages,
                 float* g_albedo, float* g_rough, float* g_sss, unsigned long long* counters, hipStream_t stream);
// Recompute entry m of a device ScatteringGPU table from device parameters
// (absorption RGB, scattering RGB in 1/mm) times `scale`.
void launch_sss_coeffs(const float* params, float scale, ScatteringGPU* table, int m, hipStream_t stream);

}  // namespace orf
