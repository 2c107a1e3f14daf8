// orf_persist.hip -- persistent batched megakernel (gfx950).
//
// One launch renders every item of a batch.  Work items are (render, 8x8 tile,
// chunk of `per` samples), numbered render by render; the item range is cut into
// 8 contiguous partitions, one per XCD (workgroup b runs on XCD b % 8), so each
// XCD's L2 works on neighbouring tiles.  A wave takes items from its XCD's
// partition with one atomic per item, then steals from the other partitions in
// turn; every wave exits when all 8 are drained.  The 8 waves of a workgroup
// share the LDS copy of the BVH's top nodes.  Each lane
// traces the chunk's samples of its pixel with the same path loop as the
// per-render megakernel and adds FP64 sums to the batch film (plain RMW when one
// chunk covers the pixel's samples, FP64 atomics otherwise).  Item fetch and the
// per-lane body are the plain megakernel's (orf_mega.h).
#include <algorithm>

#include "orf_common.h"
#include "orf_mega.h"
#include "orf_wavefront.h"

namespace orf {

constexpr int kPersistBlock = kMegaBlock;

template <unsigned kFeat>
__global__ __launch_bounds__(kPersistBlock, mega_waves(kFeat)) void k_batch_persistent(
    SceneGPU sc, const WfRender* __restrict__ renders, const PersistRender* __restrict__ pr, int nrenders,
    const int* __restrict__ remap, double* film, unsigned long long* stats, unsigned long long* qctr,
    long long total)
{
    __shared__ unsigned stack_lds[kLdsStack * kPersistBlock];
    __shared__ float4 top_lds[(kPersistTop > 0 ? kPersistTop : 1u) * kTopFloat4];
    unsigned* stk = stack_lds + threadIdx.x;
    const unsigned ntop = kPersistTop > 0 ? load_top_nodes(sc, top_lds, kPersistTop) : 0u;
    const unsigned lane = threadIdx.x & 63u;
    const int home = (int)(blockIdx.x % kXcdParts);
    unsigned long long bad = 0, npaths = 0;
    PathStats st;
    for (int q = 0; q < kXcdParts; ++q) {
        const int part = (home + q) % kXcdParts;
        const long long lo = total * part / kXcdParts, hi = total * (part + 1) / kXcdParts;
        for (;;) {
            const long long item = mega_next(qctr, part, lo, hi);
            if (item < 0) break;  // uniform across the wave
            int a = 0, b = nrenders - 1;
            while (a < b) {
                int m = (a + b + 1) >> 1;
                if (pr[m].item_begin <= item) a = m;
                else b = m - 1;
            }
            const WfRender& R = renders[a];
            const PersistRender& P = pr[a];
            const long long local = item - P.item_begin;
            const int tile = (int)(local / P.chunks), chunk = (int)(local - (long long)tile * P.chunks);
            const int x = (tile % P.tiles_x) * 8 + (int)(lane & 7), y = (tile / P.tiles_x) * 8 + (int)(lane >> 3);
            if (x < R.width && y < R.height) {  // every lane reaches the next item fetch
                const unsigned k0 = (unsigned)chunk * P.per, k1 = min(k0 + P.per, R.spp);
                mega_pixel<kFeat, kPersistTop>(sc, R.cam, R.width, R.height, x, y, remap + R.remap_offset, R.frame0, k0, k1,
                           P.chunks == 1, film + 4 * ((size_t)R.film_offset + (size_t)y * R.width + x), stk, (lds_cf4*)top_lds,
                           ntop, st, bad);
                npaths += k1 - k0;
            }
        }
    }
    atomicAdd(stats + 0, bad);
    atomicAdd(stats + 1, (unsigned long long)st.ext);
    atomicAdd(stats + 2, (unsigned long long)st.shd);
    atomicAdd(stats + 3, npaths);
}

void launch_batch_persistent(const SceneGPU& sc, const WfRender* renders, const PersistRender* pr, int nrenders,
                             const int* remap, double* film, unsigned long long* stats, unsigned long long* qctr,
                             long long total_items, int waves, hipStream_t stream)
{
    ORF_REQUIRE(nrenders > 0 && waves > 0, "persistent batch: empty launch");
    // per-partition counters are 32-bit once read back
    ORF_REQUIRE(total_items / kXcdParts < 0xF0000000LL, "persistent batch: too many work items");
    if (total_items == 0) return;
    zero_queue_counters(qctr, stream);
    const dim3 g((unsigned)mega_blocks(waves)), b(kPersistBlock);
    if (sc.features == 0u)
        hipLaunchKernelGGL(k_batch_persistent<0u>, g, b, 0, stream, sc, renders, pr, nrenders, remap, film, stats, qctr,
                           total_items);
    else if (sc.features == kFeatDielectric)
        hipLaunchKernelGGL(k_batch_persistent<kFeatDielectric>, g, b, 0, stream, sc, renders, pr, nrenders, remap, film,
                           stats, qctr, total_items);
    else
        hipLaunchKernelGGL(k_batch_persistent<kFeatAll>, g, b, 0, stream, sc, renders, pr, nrenders, remap, film,
                           stats, qctr, total_items);
    ORF_HIP_CHECK(hipGetLastError());
}

}  // namespace orf
