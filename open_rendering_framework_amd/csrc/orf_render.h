// orf_render.h -- host-side entry points of the render kernels (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "orf_scene.h"

namespace orf {

// Minimum waves per SIMD the register allocator must allow (occupancy vs
// spills) in the megakernels without / with the translucent material.
#ifndef ORF_MEGA_WAVES
#define ORF_MEGA_WAVES 8
#endif
#ifndef ORF_MEGA_WAVES_SSS
#define ORF_MEGA_WAVES_SSS 8
#endif
constexpr int mega_waves(unsigned features)
{
    return (features & kFeatTranslucent) ? ORF_MEGA_WAVES_SSS : ORF_MEGA_WAVES;
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8),
// and each XCD has its own L2.  Renumbering so that XCD x receives one contiguous
// run of logical blocks (consecutive image tiles) keeps each XCD's rays in one
// image region.  Off by default: with static dispatch the contiguous bands give
// the XCDs unequal work (the 512x512 path-tracer image ran 1.7x slower), so the
// batched persistent kernel gets locality from per-XCD queues with stealing
// instead.  A bijection on [0, nb) for any nb.
#ifndef ORF_XCD_REMAP
#define ORF_XCD_REMAP 0
#endif
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
#if ORF_XCD_REMAP
    unsigned per = nb >> 3, rem = nb & 7u, x = b & 7u, k = b >> 3;
    return x * per + (x < rem ? x : rem) + k;
#else
    (void)nb;
    return b;
#endif
}

struct FilmGPU {
    double* sums;                  // 4 per pixel: R, G, B sums and finite count
    unsigned long long* counters;  // [0] non-finite, [1] extension rays, [2] shadow rays, [3] paths
    int width, height;
    unsigned long long* qctr;      // 8 work-queue counters (device, zeroed by each launch)
};

// Waves that fill the current device at the megakernels' occupancy: CUs x 4
// SIMDs x waves, capped at the item count (each wave fetches its own items).
int persistent_waves(long long items, int waves);
// Waves per workgroup of the persistent megakernels and the number of top BVH
// nodes each workgroup keeps in LDS.
#ifndef ORF_MEGA_BLOCK_WAVES
#define ORF_MEGA_BLOCK_WAVES 8
#endif
#ifndef ORF_MEGA_TOP
#define ORF_MEGA_TOP 0
#endif
#ifndef ORF_PERSIST_TOP
#define ORF_PERSIST_TOP 200
#endif
// Workgroups of ORF_MEGA_BLOCK_WAVES waves holding `waves` waves.
inline int mega_blocks(int waves) { return (waves + ORF_MEGA_BLOCK_WAVES - 1) / ORF_MEGA_BLOCK_WAVES; }
// Zero 8 queue counters with a kernel (safe inside captured streams).
void zero_queue_counters(unsigned long long* qctr, hipStream_t stream);
// Film mean RGBA32F (n pixels) into a device buffer: RGB = sums / count, A = 1
// or depth sum / count when `depth` is given.
void launch_film_display(const double* sums, const double* depth, long long n, float4* out, hipStream_t stream);

// Megakernel: every (pixel, sample) path of frames frame0 .. frame0+spp-1.
// remap: per-render material table (device, nullptr = identity).
void launch_megakernel(const SceneGPU& sc, const CameraGPU& cam, const FilmGPU& film, const int* remap,
                       unsigned frame0, unsigned spp, hipStream_t stream);

// Depth integrator (DepthTracer): frames frame0 .. frame0+spp-1; normal colour
// into film.sums, hit distance sums into depth (one double per pixel).
void launch_depth(const SceneGPU& sc, const CameraGPU& cam, const FilmGPU& film, double* depth, unsigned frame0,
                  unsigned spp, hipStream_t stream);

// Persistent batched megakernel: one launch for every item of a batch.  Work
// items are (render, 8x8 tile, sample chunk) in render order, split into 8
// contiguous partitions, one per XCD; a wave takes items from its own XCD's
// partition, then steals from the others.
struct PersistRender {
    long long item_begin;  // first work item of the render
    int tiles_x, chunks;
    unsigned per;          // samples per chunk
    int pad;
};
struct WfRender;
// qctr: 8 zeroed u64 queue counters (device).  waves: persistent waves (grid of
// workgroups of 8 waves).
void launch_batch_persistent(const SceneGPU& sc, const WfRender* renders, const PersistRender* pr, int nrenders,
                             const int* remap, double* film, unsigned long long* stats, unsigned long long* qctr,
                             long long total_items, int waves, hipStream_t stream);

// FP64 construction of the environment's FP32 pdf/cdf tables (spec CU/env_cameras.cu).
// scratch: (h + 1) doubles.
void build_env_tables(const float4* tex, int w, int h, float* marg_pdf, float* marg_cdf, float* cond_pdf,
                      float* cond_cdf, double* scratch, hipStream_t stream);

// Conductor constants as the reference forms eta = 1 / ior (ORF/MyComplex.h:155-160).
void conductor_eta_host(const double ior_re[3], const double ior_im[3], ConductorEta& e);
void dielectric_eta_host(float n12, ConductorEta& e);

}  // namespace orf
