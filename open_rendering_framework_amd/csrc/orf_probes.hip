// orf_probes.hip -- probe pass of the translucent material (gfx950): one lane
// per probe runs probe_sample() (orf_probe_sample.h) and stores Le, w12, T12 for
// the gathers of the following render.
#include "orf_common.h"
#include "orf_probe_sample.h"
#include "orf_probes_host.h"

namespace orf {

constexpr int kProbeBlock = 64;
// 8 waves/SIMD (64 VGPRs, spilling).  Unbounded, the probe estimator takes 186
// VGPRs (2 waves/SIMD) and the 512x512, 128 spp subsurface step measured 1.7%
// slower; 4 waves/SIMD measured the same as 8.
constexpr int kProbeWaves = 8;

__global__ __launch_bounds__(kProbeBlock, kProbeWaves) void k_probe_pass(SceneGPU sc, const ScatteringGPU* __restrict__ props,
                                                            int sss, unsigned frame)
{
    __shared__ unsigned stack_lds[kLdsStack * kProbeBlock];
    int j = blockIdx.x * kProbeBlock + threadIdx.x;
    if (j >= sc.probes.count) return;
    float4 L, w, n;
    probe_sample(sc, props[sss], sc.materials[sss], j, frame, stack_lds + threadIdx.x, kProbeBlock, L, w, n);
    sc.probes.L[j] = L;
    sc.probes.w12[j] = w;
    sc.probes.sn[j] = n;
}

void launch_probe_pass(const SceneGPU& sc, const ScatteringGPU* d_props, int sss, unsigned frame, hipStream_t stream)
{
    if (sc.probes.count <= 0) return;
    ORF_REQUIRE(sc.probes.sss < 0, "probe pass: gathers must be disabled");
    unsigned grid = (unsigned)((sc.probes.count + kProbeBlock - 1) / kProbeBlock);
    hipLaunchKernelGGL(k_probe_pass, dim3(grid), dim3(kProbeBlock), 0, stream, sc, d_props, sss, frame);
    ORF_HIP_CHECK(hipGetLastError());
}

}  // namespace orf
