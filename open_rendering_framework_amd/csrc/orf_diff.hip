// orf_diff.hip -- differentiable form: primal and path-replay adjoint kernels (gfx950).
//
// Both kernels are persistent: work items are (view, 8x8 pixel tile, chunk of
// kDiffChunk samples), split into 8 contiguous partitions, one per XCD; a wave
// takes items of its XCD's partition with one atomic each, then steals from
// the others (orf_mega.h), so an expensive tile no longer holds a whole wave
// while the rest of the chip idles.  Views of a differentiable render share a
// size, so tiles are uniform.  The primal writes each chunk's FP32 sums and
// finite count per pixel; a second kernel adds a pixel's chunks in chunk order
// and writes the per-view FP32 mean image (deterministic: no atomics).  The
// adjoint computes path-replay gradient terms (each gradient vertex's share of
// the radiance still to come) and adds texel gradients with FP32 atomics (16 B
// per shading event).  Adjoint samples use frames shifted by
// `frame_shift` so they are decorrelated from the primal's.  An adjoint sample
// runs its path once, keeping each gradient vertex's terms until the radiance is
// known (diff_adjoint_sample); a path with more such vertices than that record
// holds is replayed (diff_replay), and the adjoint counts the replays that end
// on a radiance other than their first pass's (counters[1]; zero unless the two
// passes took different paths).
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "orf_common.h"
#include "orf_diff.h"
#include "orf_diff_host.h"
#include "orf_mega.h"

namespace orf {

// Waves per SIMD the register budget is sized for.  The Lambert / conductor /
// emitter primal runs at the plain megakernel's 8 waves (64 VGPRs); its adjoint
// fits 128 VGPRs without spilling (4 waves); the larger material sets keep the
// compiler's own allocation.
constexpr int diff_waves(unsigned features, bool adjoint) { return features == 0u ? (adjoint ? 4 : 8) : 1; }
// Adjoint workgroups fill a CU at that occupancy (4 SIMDs), so the CU's waves
// share one LDS copy of the texel gradients; primal workgroups hold 8 waves.
constexpr int diff_block(unsigned features, bool adjoint)
{
    return adjoint ? 256 * diff_waves(features, true) : (features == 0u ? 512 : 256);
}
// LDS a workgroup may give the texel gradients (the 64x64 albedo + roughness
// textures of the inverse-rendering workload take 64 KB); larger textures add
// straight to global memory.
constexpr size_t kDiffLdsGradBytes = 96 * 1024;

// LDS float pointer: the LDS address space gives ds_add_f32 for the gradient
// atomics (a generic pointer compiles to flat atomics, which also pass through
// the texture address unit).
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) float lds_float;
#else
typedef float lds_float;
#endif

// Texel gradients go to FP32 arrays through atomics: the workgroup's LDS copy
// when the launch gives one (flushed to global memory once at the end), else
// the global arrays; the translucent coefficients' six gradients are summed per
// lane and added once per lane.
struct AtomicSink {
    float* g_albedo;
    float* g_rough;
    float* g_sss;  // d/d absorption RGB, d/d scattering RGB
    f3 ga, gs;
    lds_float* l_albedo;  // LDS copies (null: global atomics)
    lds_float* l_rough;
    __device__ void albedo(int texel, int c, float v)
    {
        if (l_albedo) {
            atomicAdd(l_albedo + 3 * texel + c, v);
            // keeps the two arms apart: merged, they become one flat atomic
            __asm__ volatile("" ::: "memory");
        } else if (g_albedo) {
            atomicAdd(g_albedo + 3 * (size_t)texel + c, v);
        }
    }
    __device__ void rough(int texel, float v)
    {
        if (l_rough) {
            atomicAdd(l_rough + texel, v);
            __asm__ volatile("" ::: "memory");
        } else if (g_rough) {
            atomicAdd(g_rough + texel, v);
        }
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

template <unsigned kFeat, bool kAdjoint>
__global__ __launch_bounds__(diff_block(kFeat, kAdjoint), diff_waves(kFeat, kAdjoint)) void k_diff(
    SceneGPU sc, DiffParams dp, const WfRender* __restrict__ views, int tiles_x, int tiles_per_view, int chunks,
    long long items, unsigned frame_shift, float4* partial, long long npix, const float* dimages, float* g_albedo,
    float* g_rough, float* g_sss, unsigned long long* counters, unsigned long long* qctr, int lds_grad)
{
    constexpr int kDiffBlock = diff_block(kFeat, kAdjoint);
    constexpr bool adjoint = kAdjoint;
    __shared__ unsigned stack_lds[kLdsStack * kDiffBlock];
    unsigned* stk = stack_lds + threadIdx.x;
    const unsigned lane = threadIdx.x & 63u;
    const int home = (int)(blockIdx.x % kXcdParts);
    unsigned bad = 0, mismatch = 0;
    AtomicSink sink{g_albedo, g_rough, g_sss, mk3(0.0f), mk3(0.0f), nullptr, nullptr};
    // the workgroup's LDS copy of the texel gradients (dynamic LDS, zeroed here)
    extern __shared__ float grad_lds[];
    const int texels = dp.tw * dp.th;
    const int n_albedo = g_albedo ? 3 * texels : 0, n_rough = g_rough ? texels : 0;
    if (lds_grad) {
        sink.l_albedo = g_albedo ? (lds_float*)grad_lds : nullptr;
        sink.l_rough = g_rough ? (lds_float*)grad_lds + n_albedo : nullptr;
        for (int i = threadIdx.x; i < n_albedo + n_rough; i += kDiffBlock) grad_lds[i] = 0.0f;
        __syncthreads();
    }
    for (int q = 0; q < kXcdParts; ++q) {
        const int part = (home + q) % kXcdParts;
        const long long lo = items * part / kXcdParts, hi = items * (part + 1) / kXcdParts;
        for (;;) {
            const long long item = mega_next(qctr, part, lo, hi);
            if (item < 0) break;  // uniform across the wave
            const long long vt = item / chunks;
            const int chunk = (int)(item - vt * chunks);
            const int view = (int)(vt / tiles_per_view), tile = (int)(vt - (long long)view * tiles_per_view);
            const WfRender& R = views[view];
            const int x = (tile % tiles_x) * 8 + (int)(lane & 7), y = (tile / tiles_x) * 8 + (int)(lane >> 3);
            if (x >= R.width || y >= R.height) continue;  // every lane reaches the next item fetch
            const size_t px = ((size_t)view * R.height + y) * R.width + x;
            const unsigned k0 = (unsigned)chunk * kDiffChunk, k1 = min(k0 + kDiffChunk, R.spp);
            if constexpr (!kAdjoint) {
                f3 acc = mk3(0.0f);
                unsigned cnt = 0;
                for (unsigned s = k0; s < k1; ++s) {
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
                partial[(size_t)chunk * (size_t)npix + px] = make_float4(acc.x, acc.y, acc.z, (float)cnt);
            } else {
                const float sc_inv = 1.0f / (float)max(1u, R.spp);
                const f3 delta = mk3(dimages[3 * px], dimages[3 * px + 1], dimages[3 * px + 2]) * sc_inv;
                if (delta.x == 0.0f && delta.y == 0.0f && delta.z == 0.0f) continue;
                for (unsigned s = k0; s < k1; ++s) {
                    f3 dir;
                    uint32_t s0;
                    uint32_t rng = camera_ray(R.cam, R.width, R.height, x, y, R.frame0 + frame_shift + s, dir, &s0);
                    bool spill;
                    f3 L = diff_adjoint_sample<kFeat>(sc, dp, R.cam.eye, dir, rng, seed64(s0), delta, sink, stk,
                                                      kDiffBlock, spill);
                    if (!finite3(L)) {
                        bad++;
                        continue;
                    }
                    if (spill) {  // more gradient vertices than the record holds: replay the path
                        const f3 Lr = diff_replay<kFeat>(sc, dp, R.cam.eye, dir, rng, seed64(s0), L, delta, sink,
                                                         stk, kDiffBlock);
                        if (!(Lr.x == L.x && Lr.y == L.y && Lr.z == L.z)) mismatch++;
                    }
                }
            }
        }
    }
    if (adjoint) sink.flush();
    if (lds_grad) {  // every wave of the workgroup has left its item loop
        __syncthreads();
        for (int i = threadIdx.x; i < n_albedo; i += kDiffBlock) {
            const float v = grad_lds[i];
            if (v != 0.0f) atomicAdd(g_albedo + i, v);
        }
        for (int i = threadIdx.x; i < n_rough; i += kDiffBlock) {
            const float v = grad_lds[n_albedo + i];
            if (v != 0.0f) atomicAdd(g_rough + i, v);
        }
    }
    if (bad) atomicAdd(counters, (unsigned long long)bad);
    if (mismatch) atomicAdd(counters + 1, (unsigned long long)mismatch);
}

// Mean image of every pixel from its chunks' sums, added in chunk order.
__global__ __launch_bounds__(256) void k_diff_image(const float4* __restrict__ partial, int chunks, long long npix,
                                                    float* __restrict__ images)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    f3 acc = mk3(0.0f);
    float cnt = 0.0f;
    for (int c = 0; c < chunks; ++c) {
        const float4 p = partial[(size_t)c * (size_t)npix + i];
        acc += mk3(p.x, p.y, p.z);
        cnt += p.w;
    }
    const float inv = cnt > 0.0f ? 1.0f / cnt : 0.0f;
    images[3 * i + 0] = acc.x * inv;
    images[3 * i + 1] = acc.y * inv;
    images[3 * i + 2] = acc.z * inv;
}

void launch_diff(const SceneGPU& sc, const DiffParams& dp, const WfRender* d_views, int nviews, int width, int height,
                 unsigned spp, bool adjoint, unsigned frame_shift, float* images, float4* partial, const float* dimages,
                 float* g_albedo, float* g_rough, float* g_sss, unsigned long long* counters,
                 unsigned long long* qctr, hipStream_t stream)
{
    const int tiles_x = (width + 7) / 8, tiles_y = (height + 7) / 8;
    const int chunks = (int)((spp + kDiffChunk - 1) / kDiffChunk);
    const long long items = (long long)tiles_x * tiles_y * nviews * chunks;
    const long long npix = (long long)width * height * nviews;
    ORF_REQUIRE(items > 0 && items / kXcdParts < 0xF0000000LL && npix < 0x7fffffffLL * 256LL, "diff launch: bad size");
    ORF_REQUIRE(adjoint || partial, "diff launch: no partial-sum buffer");
    zero_queue_counters(qctr, stream);
    zero_queue_counters(counters, stream);  // this launch's statistics (8 entries)
    const int waves = diff_waves(sc.features, adjoint), bsize = diff_block(sc.features, adjoint), wpb = bsize / 64;
    const dim3 grid((unsigned)((persistent_waves(items, waves) + wpb - 1) / wpb)), block((unsigned)bsize);
    // texel gradients in LDS when they fit beside the traversal stacks
    const size_t grad_bytes =
        adjoint ? ((g_albedo ? 3u : 0u) + (g_rough ? 1u : 0u)) * (size_t)dp.tw * (size_t)dp.th * sizeof(float) : 0;
    int dev = 0, lds_max = 0;
    ORF_HIP_CHECK(hipGetDevice(&dev));
    ORF_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t stack_bytes = (size_t)kLdsStack * bsize * sizeof(unsigned);
    const char* force_global = std::getenv("ORF_DIFF_GLOBAL_GRAD_ATOMICS");
    const bool lds = grad_bytes > 0 && grad_bytes <= kDiffLdsGradBytes &&
                     stack_bytes + grad_bytes <= (size_t)lds_max && !(force_global && force_global[0] == '1');
    // the scene's material programs only, as the render kernels pick them
    auto pick = [&](auto adj) {
        constexpr bool kA = decltype(adj)::value;
        return sc.features == 0u                 ? k_diff<0u, kA>
               : sc.features == kFeatDielectric ? k_diff<kFeatDielectric, kA>
                                                 : k_diff<kFeatAll, kA>;
    };
    auto kernel = adjoint ? pick(std::true_type{}) : pick(std::false_type{});
    hipLaunchKernelGGL(kernel, grid, block, lds ? grad_bytes : 0, stream, sc, dp, d_views, tiles_x, tiles_x * tiles_y,
                       chunks, items, frame_shift, partial, npix, dimages, g_albedo, g_rough, g_sss, counters, qctr,
                       lds ? 1 : 0);
    ORF_HIP_CHECK(hipGetLastError());
    if (!adjoint) {
        hipLaunchKernelGGL(k_diff_image, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, stream, partial, chunks,
                           npix, images);
        ORF_HIP_CHECK(hipGetLastError());
    }
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
