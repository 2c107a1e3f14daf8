// orf_diff_host.h -- host entry of the differentiable kernels (internal).
#pragma once

#include "orf_diff.h"
#include "orf_wavefront.h"

namespace orf {

// One launch over `nviews` views of width x height: primal (images) or adjoint
// (dimages -> g_albedo, g_rough).  counters[0] receives dropped non-finite samples.
void launch_diff(const SceneGPU& sc, const DiffParams& dp, const WfRender* d_views, int nviews, int width,
                 int height, bool adjoint, unsigned frame_shift, float* images, const float* dimages,
                 float* g_albedo, float* g_rough, unsigned long long* counters, hipStream_t stream);

}  // namespace orf
