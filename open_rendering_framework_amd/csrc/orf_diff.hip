// orf_diff.hip -- differentiable form: primal and path-replay adjoint kernels (gfx950).
//
// Both kernels map one wave64 workgroup to an 8x8 pixel tile of one view and run
// every sample of a pixel in one lane (views of a differentiable render share a
// size, so tiles are uniform).  The primal writes the per-view FP32 mean image;
// the adjoint replays each sample twice (pass 1: radiance L; pass 2: gradient
// terms against the radiance still to come) and adds texel gradients with FP32
// atomics (16 B per shading event).  Adjoint samples use frames shifted by
// `frame_shift` so they are decorrelated from the primal's.
#include <algorithm>

#include "orf_common.h"
#include "orf_diff.h"
#include "orf_diff_host.h"

namespace orf {

constexpr int kDiffBlock = 64;

// Waves per SIMD the register budget is sized for: the Lambert / conductor /
// emitter form fits 128 VGPRs without spilling (4 waves); the larger material
// sets keep the compiler's own allocation.
constexpr int diff_waves(unsigned features) { return features == 0u ? 4 : 1; }

// Texel gradients land in global FP32 arrays through atomics; the translucent
// coefficients' six gradients are summed per lane and added once per lane.
struct AtomicSink {
    float* g_albedo;
    float* g_rough;
    float* g_sss;  // d/d absorption RGB, d/d scattering RGB
    f3 ga, gs;
    __device__ void albedo(int texel, int c, float v)
    {
        if (g_albedo) atomicAdd(g_albedo + 3 * (size_t)texel + c, v);
    }
    __device__ void rough(int texel, float v)
    {
        if (g_rough) atomicAdd(g_rough + texel, v);
    }
    __device__ void sss(f3 a, f3 s)
    {
        ga += a;
        gs += s;
    }
    __device__ void flush()
    {
        if (!g_sss) return;
        const float v[6] = {ga.x, ga.y, ga.z, gs.x, gs.y, gs.z};
        for (int k = 0; k < 6; ++k)
            if (v[k] != 0.0f) atomicAdd(g_sss + k, v[k]);
    }
};

template <unsigned kFeat>
__global__ __launch_bounds__(kDiffBlock, diff_waves(kFeat)) void k_diff(SceneGPU sc, DiffParams dp, const WfRender* __restrict__ views,
                                                     int tiles_x, int tiles_per_view, int adjoint,
                                                     unsigned frame_shift, float* images, const float* dimages,
                                                     float* g_albedo, float* g_rough, float* g_sss,
                                                     unsigned long long* counters)
{
    __shared__ unsigned stack_lds[kLdsStack * kDiffBlock];
    unsigned* stk = stack_lds + threadIdx.x;
    int view = blockIdx.x / tiles_per_view;
    int tile = blockIdx.x % tiles_per_view;
    const WfRender& R = views[view];
    int x = (tile % tiles_x) * 8 + (threadIdx.x & 7);
    int y = (tile / tiles_x) * 8 + (threadIdx.x >> 3);
    if (x >= R.width || y >= R.height) return;
    size_t px = ((size_t)view * R.height + y) * R.width + x;
    unsigned bad = 0;
    if (!adjoint) {
        f3 acc = mk3(0.0f);
        unsigned cnt = 0;
        for (unsigned s = 0; s < R.spp; ++s) {
            f3 dir;
            uint32_t s0;
            uint32_t rng = camera_ray(R.cam, R.width, R.height, x, y, R.frame0 + s, dir, &s0);
            f3 L = diff_path<kFeat>(sc, dp, R.cam.eye, dir, rng, seed64(s0), stk, kDiffBlock);
            if (finite3(L)) {
                acc += L;
                cnt++;
            } else {
                bad++;
            }
        }
        float inv = cnt ? 1.0f / (float)cnt : 0.0f;
        images[3 * px + 0] = acc.x * inv;
        images[3 * px + 1] = acc.y * inv;
        images[3 * px + 2] = acc.z * inv;
    } else {
        AtomicSink sink{g_albedo, g_rough, g_sss, mk3(0.0f), mk3(0.0f)};
        float sc_inv = 1.0f / (float)max(1u, R.spp);
        f3 delta = mk3(dimages[3 * px], dimages[3 * px + 1], dimages[3 * px + 2]) * sc_inv;
        if (delta.x == 0.0f && delta.y == 0.0f && delta.z == 0.0f) return;
        for (unsigned s = 0; s < R.spp; ++s) {
            f3 dir;
            uint32_t s0;
            uint32_t rng = camera_ray(R.cam, R.width, R.height, x, y, R.frame0 + frame_shift + s, dir, &s0);
            f3 L = diff_path<kFeat>(sc, dp, R.cam.eye, dir, rng, seed64(s0), stk, kDiffBlock);
            if (!finite3(L)) {
                bad++;
                continue;
            }
            diff_replay<kFeat>(sc, dp, R.cam.eye, dir, rng, seed64(s0), L, delta, sink, stk, kDiffBlock);
        }
        sink.flush();
    }
    if (bad) atomicAdd(counters, (unsigned long long)bad);
}

void launch_diff(const SceneGPU& sc, const DiffParams& dp, const WfRender* d_views, int nviews, int width,
                 int height, bool adjoint, unsigned frame_shift, float* images, const float* dimages,
                 float* g_albedo, float* g_rough, float* g_sss, unsigned long long* counters, hipStream_t stream)
{
    int tiles_x = (width + 7) / 8, tiles_y = (height + 7) / 8;
    int64_t blocks = (int64_t)tiles_x * tiles_y * nviews;
    ORF_REQUIRE(blocks > 0 && blocks < 0x7fffffffLL, "diff launch: bad size");
    // the scene's material programs only, as the render kernels pick them
    auto kernel = sc.features == 0u ? k_diff<0u> : sc.features == kFeatDielectric ? k_diff<kFeatDielectric>
                                                                                   : k_diff<kFeatAll>;
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(kDiffBlock), 0, stream, sc, dp, d_views, tiles_x,
                       tiles_x * tiles_y, adjoint ? 1 : 0, frame_shift, images, dimages, g_albedo, g_rough, g_sss,
                       counters);
    ORF_HIP_CHECK(hipGetLastError());
}

}  // namespace orf

namespace orf {

// The coefficients of material `m` from device parameters (absorption RGB,
// scattering RGB, unscaled), the rest of its record kept: one lane.
__global__ void k_sss_coeffs(const float* __restrict__ params, float scale, ScatteringGPU* table, int m)
{
    if (threadIdx.x != 0) return;
    const ScatteringGPU t = table[m];
    const f3 sa = mk3(params[0], params[1], params[2]) * scale;
    const f3 ss = mk3(params[3], params[4], params[5]) * scale;
    table[m] = sss_coefficients(sa, ss, t.meancosine, t.relative_ior, t.dipole_model);
}

void launch_sss_coeffs(const float* params, float scale, ScatteringGPU* table, int m, hipStream_t stream)
{
    hipLaunchKernelGGL(k_sss_coeffs, dim3(1), dim3(64), 0, stream, params, scale, table, m);
    ORF_HIP_CHECK(hipGetLastError());
}

}  // namespace orf
