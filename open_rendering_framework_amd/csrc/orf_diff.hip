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

__device__ f3 diff_path(const SceneGPU& sc, const DiffParams& dp, f3 o, f3 d, uint32_t rng, unsigned* stk)
{
    PathState ps;
    path_begin(ps, o, d, rng);
    for (;;) {
        Hit h;
        if (!trace<false>(sc, ps.o, ps.d, sc.scene_epsilon, kRayMax, h, stk, kDiffBlock)) {
            shade_miss(sc, ps);
            break;
        }
        ShadowRay sr;
        DiffVertex dv;
        bool more = diff_vertex(sc, dp, ps, h, sr, dv);
        if (sr.valid) {
            Hit hs;
            if (!trace<true>(sc, sr.o, sr.d, sc.scene_epsilon, sr.tmax, hs, stk, kDiffBlock)) ps.L += sr.contrib;
        }
        if (!more) break;
    }
    return ps.L;
}

// Pass 2: replay and scatter gradients; delta = dLoss/dL of this sample.
__device__ void diff_replay(const SceneGPU& sc, const DiffParams& dp, f3 o, f3 d, uint32_t rng, f3 L_total,
                            f3 delta, float* g_albedo, float* g_rough, unsigned* stk)
{
    PathState ps;
    path_begin(ps, o, d, rng);
    for (;;) {
        Hit h;
        if (!trace<false>(sc, ps.o, ps.d, sc.scene_epsilon, kRayMax, h, stk, kDiffBlock)) break;
        f3 rem_before = L_total - ps.L;  // radiance from this vertex on
        ShadowRay sr;
        DiffVertex dv;
        bool more = diff_vertex(sc, dp, ps, h, sr, dv);
        if (sr.valid) {
            Hit hs;
            if (!trace<true>(sc, sr.o, sr.d, sc.scene_epsilon, sr.tmax, hs, stk, kDiffBlock)) ps.L += sr.contrib;
        }
        if (dv.kind == 1 && g_albedo) {
            float* g = g_albedo + 3 * (size_t)dv.texel;
            if (dv.rho.x != 0.0f) atomicAdd(g + 0, delta.x * rem_before.x / dv.rho.x);
            if (dv.rho.y != 0.0f) atomicAdd(g + 1, delta.y * rem_before.y / dv.rho.y);
            if (dv.rho.z != 0.0f) atomicAdd(g + 2, delta.z * rem_before.z / dv.rho.z);
        } else if (dv.kind == 2 && g_rough) {
            f3 rem_after = L_total - ps.L;
            atomicAdd(g_rough + dv.texel, dot(delta, rem_after) * dv.dlogw);
        }
        if (!more) break;
    }
}

__global__ __launch_bounds__(kDiffBlock) void k_diff(SceneGPU sc, DiffParams dp, const WfRender* __restrict__ views,
                                                     int tiles_x, int tiles_per_view, int adjoint,
                                                     unsigned frame_shift, float* images, const float* dimages,
                                                     float* g_albedo, float* g_rough,
                                                     unsigned long long* counters)
{
    __shared__ unsigned stack_lds[kStackDepth * kDiffBlock];
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
            uint32_t rng = camera_ray(R.cam, R.width, R.height, x, y, R.frame0 + s, dir);
            f3 L = diff_path(sc, dp, R.cam.eye, dir, rng, stk);
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
        float sc_inv = 1.0f / (float)max(1u, R.spp);
        f3 delta = mk3(dimages[3 * px], dimages[3 * px + 1], dimages[3 * px + 2]) * sc_inv;
        if (delta.x == 0.0f && delta.y == 0.0f && delta.z == 0.0f) return;
        for (unsigned s = 0; s < R.spp; ++s) {
            f3 dir;
            uint32_t rng = camera_ray(R.cam, R.width, R.height, x, y, R.frame0 + frame_shift + s, dir);
            f3 L = diff_path(sc, dp, R.cam.eye, dir, rng, stk);
            if (!finite3(L)) {
                bad++;
                continue;
            }
            diff_replay(sc, dp, R.cam.eye, dir, rng, L, delta, g_albedo, g_rough, stk);
        }
    }
    if (bad) atomicAdd(counters, (unsigned long long)bad);
}

void launch_diff(const SceneGPU& sc, const DiffParams& dp, const WfRender* d_views, int nviews, int width,
                 int height, bool adjoint, unsigned frame_shift, float* images, const float* dimages,
                 float* g_albedo, float* g_rough, unsigned long long* counters, hipStream_t stream)
{
    int tiles_x = (width + 7) / 8, tiles_y = (height + 7) / 8;
    int64_t blocks = (int64_t)tiles_x * tiles_y * nviews;
    ORF_REQUIRE(blocks > 0 && blocks < 0x7fffffffLL, "diff launch: bad size");
    hipLaunchKernelGGL(k_diff, dim3((unsigned)blocks), dim3(kDiffBlock), 0, stream, sc, dp, d_views, tiles_x,
                       tiles_x * tiles_y, adjoint ? 1 : 0, frame_shift, images, dimages, g_albedo, g_rough,
                       counters);
    ORF_HIP_CHECK(hipGetLastError());
}

}  // namespace orf
