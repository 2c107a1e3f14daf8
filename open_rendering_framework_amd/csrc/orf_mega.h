// orf_mega.h -- per-lane body and work fetch shared by the persistent megakernels (device).
//
// Work item = (8x8 pixel tile, chunk of up to kMegaChunk samples) of one render.
// A launch's items are cut into kXcdParts contiguous partitions; the waves of
// workgroup b start on partition b % kXcdParts -- the XCD they run on, as
// workgroups are dealt to the 8 XCDs round-robin -- and steal from the other
// partitions once their own is drained, so neighbouring tiles share an L2 while
// no XCD idles on a cheap image region.  Each wave fetches its own items; the
// waves of a workgroup share only the LDS copy of the BVH's top nodes, which
// they load together before the first item.
#pragma once

#include "orf_path.h"
#include "orf_render.h"

namespace orf {

// 8 wave64s per workgroup: at 8 waves per SIMD a CU holds 4 such workgroups, each
// with an 8-entry LDS stack column per lane (16 KB) and up to 200 top BVH nodes
// (112 B each, 22.4 KB) -- 154 KB of the CU's 160 KB.  On the 327,684-triangle
// scene the first 200 breadth-first nodes take about 45% of all node visits.
constexpr int kMegaWavesPerBlock = ORF_MEGA_BLOCK_WAVES;
constexpr int kMegaBlock = 64 * kMegaWavesPerBlock;
// Top BVH nodes a workgroup keeps in LDS, per kernel (0: every node from global
// memory).  Measured: the batched persistent kernel gains 5% with 200 nodes; the
// plain 512x512 launch loses 5-7% with 21, 64 or 200 (the LDS and global fetch
// arms of the traversal add VALU work that outweighs the vector-memory
// instructions saved), and the subsurface scene loses 14%.
constexpr unsigned kMegaTop = ORF_MEGA_TOP;
constexpr unsigned kPersistTop = ORF_PERSIST_TOP;
static_assert(kMegaTop <= (unsigned)kBvhTopOrdered && kPersistTop <= (unsigned)kBvhTopOrdered,
              "LDS top nodes must lie in the breadth-first prefix");
constexpr int kXcdParts = 8;
// Most samples per work item.  A lane runs its item's samples back to back, so
// longer runs leave fewer lanes idle while a wave's longest run finishes; the
// per-render launch shrinks the chunk until every wave has kMegaItemsPerWave
// items (512x512 at 64 spp gets 2), the batched persistent launch, with
// hundreds of items per wave, takes up to 16 (1,519 -> 1,628 Mpaths/s on the
// 64-render batch against 4; 32 measured the same, 64 slower).
#ifndef ORF_MEGA_CHUNK
#define ORF_MEGA_CHUNK 16
#endif
constexpr unsigned kMegaChunk = ORF_MEGA_CHUNK;
#ifndef ORF_MEGA_ITEMS_PER_WAVE
#define ORF_MEGA_ITEMS_PER_WAVE 16
#endif
constexpr long long kMegaItemsPerWave = ORF_MEGA_ITEMS_PER_WAVE;  // fewest items per wave before chunks shrink

// Next item of partition `part` ([lo, hi)) for the whole wave, or -1 when drained.
__device__ __forceinline__ long long mega_next(unsigned long long* qctr, int part, long long lo, long long hi)
{
    unsigned got = 0;
    if ((threadIdx.x & 63u) == 0u) got = (unsigned)atomicAdd(qctr + part, 1ull);
    got = __shfl(got, 0);
    long long item = lo + (long long)got;
    return item < hi ? item : -1;
}

// Samples frame0 + k, k in [k0, k1), of pixel (x, y): FP64 sums of the finite
// radiances into px (plain read-modify-write when this item is the pixel's only
// writer in the launch, FP64 atomics otherwise).  Pixels outside the camera
// patch (CU/path_tracer.cu:36-43) record zero radiance.
//
// The lane runs its samples as one loop of bounces (extension ray, shading,
// shadow ray); a finished sample starts the next at once, so a wave waits for
// its longest run of samples, not for its longest path at every sample.  The
// per-sample order of operations is path_radiance()'s, so is the radiance.
template <unsigned kFeat, unsigned kTopN>
__device__ __forceinline__ void mega_pixel(const SceneGPU& sc, const CameraGPU& cam, int w, int h, int x, int y,
                                           const int* remap, unsigned frame0, unsigned k0, unsigned k1, bool owner,
                                           double* px, unsigned* stk, lds_cf4* top, unsigned ntop, PathStats& st,
                                           unsigned long long& bad)
{
    const bool inside = !((unsigned)x < cam.patch_origin[0] || (unsigned)x > cam.patch_origin[0] + cam.patch_dims[0] ||
                          (unsigned)y < cam.patch_origin[1] || (unsigned)y > cam.patch_origin[1] + cam.patch_dims[1]);
    double sr = 0.0, sg = 0.0, sb = 0.0;
    unsigned cnt = 0;
    unsigned k = k0;
    if (!inside) {  // zero radiance for every sample
        cnt = k1 - k0;
        k = k1;
    }
    PathState ps;
    bool live = false;
    for (;;) {
        if (!live) {
            if (k >= k1) break;
            f3 dir;
            uint32_t s0;
            uint32_t rng = camera_ray(cam, w, h, x, y, frame0 + k, dir, &s0);
            path_begin(ps, cam.eye, dir, rng, seed64(s0));
            live = true;
        }
        Hit hit;
        st.ext++;
        bool more = false;
        bool found;
        if constexpr (kTopN > 0)
            found = trace_top<false, kFeat == kFeatAll, (kFeat != 0u)>(sc, ps.o, ps.d, sc.scene_epsilon, kRayMax, hit,
                                                                       stk, kMegaBlock, top, ntop);
        else
            found = trace<false, kFeat == kFeatAll, (kFeat != 0u)>(sc, ps.o, ps.d, sc.scene_epsilon, kRayMax, hit, stk,
                                                                   kMegaBlock);
        if (!found) {
            shade_miss(sc, ps);
        } else {
            ShadowRay shadow;
            more = shade_vertex<kFeat>(sc, remap, ps, hit, shadow, stk, kMegaBlock);
            if (shadow.valid) {
                Hit hs;
                st.shd++;
                bool occluded;
                if constexpr (kTopN > 0)
                    occluded = trace_top<true, kFeat == kFeatAll, (kFeat != 0u)>(
                        sc, shadow.o, shadow.d, sc.scene_epsilon, shadow.tmax, hs, stk, kMegaBlock, top, ntop);
                else
                    occluded = trace<true, kFeat == kFeatAll, (kFeat != 0u)>(sc, shadow.o, shadow.d, sc.scene_epsilon,
                                                                             shadow.tmax, hs, stk, kMegaBlock);
                if (!occluded)
                    shadow_unoccluded(ps, shadow);
            }
        }
        if (!more) {
            f3 L = ps.L;
            if (finite3(L)) {
                sr += (double)L.x;
                sg += (double)L.y;
                sb += (double)L.z;
                cnt++;
            } else {
                bad++;
            }
            ++k;
            live = false;
        }
    }
    if (owner) {
        px[0] += sr;
        px[1] += sg;
        px[2] += sb;
        px[3] += (double)cnt;
    } else {
        atomicAdd(px + 0, sr);
        atomicAdd(px + 1, sg);
        atomicAdd(px + 2, sb);
        atomicAdd(px + 3, (double)cnt);
    }
}

}  // namespace orf
