// orf_render.hip -- plain megakernel path tracer and environment CDF build (gfx950).
//
// Megakernel: persistent wave64 workgroups take (8x8 pixel tile, chunk of up
// to kMegaChunk samples, sized per launch) work items from per-XCD queues
// (orf_mega.h); each lane walks its pixel's samples of the chunk and adds FP64
// sums of its finite samples to the film (plain RMW when one chunk covers all
// samples, else FP64 atomics).
#include <algorithm>

#include "orf_common.h"
#include "orf_mega.h"

namespace orf {

template <unsigned kFeat>
__global__ __launch_bounds__(kMegaBlock, mega_waves(kFeat)) void k_render_mega(SceneGPU sc, CameraGPU cam, FilmGPU film,
                                                                             const int* remap, unsigned frame0,
                                                                             unsigned spp, unsigned per, int tiles_x,
                                                                             int chunks, long long items)
{
    __shared__ unsigned stack_lds[kLdsStack * kMegaBlock];
    __shared__ float4 top_lds[(kMegaTop > 0 ? kMegaTop : 1u) * kTopFloat4];
    unsigned* stk = stack_lds + threadIdx.x;
    const unsigned ntop = kMegaTop > 0 ? load_top_nodes(sc, top_lds, kMegaTop) : 0u;
    const unsigned lane = threadIdx.x & 63u;
    const int home = (int)(blockIdx.x % kXcdParts);
    unsigned long long bad = 0, npaths = 0;
    PathStats st;
    for (int q = 0; q < kXcdParts; ++q) {
        const int part = (home + q) % kXcdParts;
        const long long lo = items * part / kXcdParts, hi = items * (part + 1) / kXcdParts;
        for (;;) {
            const long long item = mega_next(film.qctr, part, lo, hi);
            if (item < 0) break;  // uniform across the wave
            const int tile = (int)(item / chunks), chunk = (int)(item - (long long)tile * chunks);
            const int x = (tile % tiles_x) * 8 + (int)(lane & 7), y = (tile / tiles_x) * 8 + (int)(lane >> 3);
            if (x < film.width && y < film.height) {
                const unsigned k0 = (unsigned)chunk * per, k1 = min(k0 + per, spp);
                mega_pixel<kFeat, kMegaTop>(sc, cam, film.width, film.height, x, y, remap, frame0, k0, k1, chunks == 1,
                                            film.sums + 4 * ((size_t)y * film.width + x), stk, (lds_cf4*)top_lds, ntop,
                                            st, bad);
                npaths += k1 - k0;
            }
        }
    }
    if (bad) atomicAdd(film.counters + 0, bad);
    atomicAdd(film.counters + 1, (unsigned long long)st.ext);
    atomicAdd(film.counters + 2, (unsigned long long)st.shd);
    atomicAdd(film.counters + 3, npaths);
}

// Film mean as RGBA32F into a caller's device buffer (a display buffer): RGB =
// FP64 sums / count, A = 1, or the mean hit distance with the depth sums.
__global__ __launch_bounds__(256) void k_film_display(const double* __restrict__ sums,
                                                      const double* __restrict__ depth, long long n,
                                                      float4* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double cnt = sums[4 * i + 3];
    const double r = cnt > 0 ? sums[4 * i + 0] / cnt : 0.0;
    const double g = cnt > 0 ? sums[4 * i + 1] / cnt : 0.0;
    const double b = cnt > 0 ? sums[4 * i + 2] / cnt : 0.0;
    const double a = depth ? (cnt > 0 ? depth[i] / cnt : 0.0) : 1.0;
    out[i] = make_float4((float)r, (float)g, (float)b, (float)a);
}

void launch_film_display(const double* sums, const double* depth, long long n, float4* out, hipStream_t stream)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_film_display, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, sums, depth, n, out);
    ORF_HIP_CHECK(hipGetLastError());
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

void zero_queue_counters(unsigned long long* qctr, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, qctr, kXcdParts);
    ORF_HIP_CHECK(hipGetLastError());
}

int persistent_waves(long long items, int waves)
{
    static int cus[64] = {0};
    int dev = 0;
    ORF_HIP_CHECK(hipGetDevice(&dev));
    int& n = cus[dev & 63];
    if (n == 0) ORF_HIP_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    long long g = (long long)n * 4 * waves;
    return (int)std::max<long long>(1, std::min<long long>(g, items));
}

void launch_megakernel(const SceneGPU& sc, const CameraGPU& cam, const FilmGPU& film, const int* remap,
                       unsigned frame0, unsigned spp, hipStream_t stream)
{
    if (spp == 0 || film.width <= 0 || film.height <= 0) return;
    ORF_REQUIRE(film.qctr, "launch: film has no queue counters");
    const int tiles_x = (film.width + 7) / 8, tiles_y = (film.height + 7) / 8;
    const int waves = mega_waves(sc.features);
    // Samples per item: the largest of kMegaChunk, /2, ... that still gives every
    // wave of a full persistent grid kMegaItemsPerWave items (twice as many with
    // translucent materials, whose gathers make tiles differ most in cost).
    // Larger chunks shorten the waves' idle lanes per path but leave the chip
    // idle behind the last expensive items.  Measured: the 512x512, 64 spp path
    // scene is best at 16 items per wave (2 samples per item; 32 items, 1 sample:
    // -2.3%); the 512x512, 128 spp subsurface scene at 32 (2 samples per item;
    // +2.9% over 16 items of 4 samples, the same as 64 items of 1 sample).
    const long long full = (long long)persistent_waves(1LL << 40, waves);
    const long long per_wave = (sc.features & kFeatTranslucent) ? 2 * kMegaItemsPerWave : kMegaItemsPerWave;
    unsigned per = kMegaChunk;
    while (per > 1 && (long long)tiles_x * tiles_y * ((spp + per - 1) / per) < per_wave * full) per >>= 1;
    per = std::min(per, spp);
    const int chunks = (int)((spp + per - 1) / per);
    const long long items = (long long)tiles_x * tiles_y * chunks;
    ORF_REQUIRE(items / kXcdParts < 0xF0000000LL, "launch: too many work items");
    zero_queue_counters(film.qctr, stream);
    const dim3 grid((unsigned)mega_blocks(persistent_waves(items, waves))), block(kMegaBlock);
    if (sc.features == 0u)
        hipLaunchKernelGGL(k_render_mega<0u>, grid, block, 0, stream, sc, cam, film, remap, frame0, spp, per, tiles_x,
                           chunks, items);
    else if (sc.features == kFeatDielectric)
        hipLaunchKernelGGL(k_render_mega<kFeatDielectric>, grid, block, 0, stream, sc, cam, film, remap, frame0, spp,
                           per, tiles_x, chunks, items);
    else
        hipLaunchKernelGGL(k_render_mega<kFeatAll>, grid, block, 0, stream, sc, cam, film, remap, frame0, spp, per,
                           tiles_x, chunks, items);
    ORF_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ env CDF

// One workgroup per row: luminance * sin(theta) in FP64 (NTSC weights of
// CU/env_cameras.cu:18-21 at texel centres, :23-29), row prefix sums, conditional
// pdf/cdf (:45-58); the row mean goes to scratch[row].  Zero rows get a uniform
// conditional distribution (their marginal pdf is 0, so they are never sampled).
constexpr int kEnvBlock = 256;

__global__ __launch_bounds__(kEnvBlock) void k_env_rows(const float4* __restrict__ tex, int w, int h,
                                                        float* __restrict__ cond_pdf, float* __restrict__ cond_cdf,
                                                        double* __restrict__ row_mean)
{
    __shared__ double part[kEnvBlock];
    int row = blockIdx.x;
    double theta = (1.0 - ((double)row + 0.5) / (double)h) * 3.14159265358979323846;
    double st = sin(theta);
    int per = (w + kEnvBlock - 1) / kEnvBlock;
    int i0 = threadIdx.x * per, i1 = min(w, i0 + per);
    double local = 0.0;
    for (int i = i0; i < i1; ++i) {
        float4 t = tex[(size_t)row * w + i];
        local += ((double)t.x * 0.2989 + (double)t.y * 0.5866 + (double)t.z * 0.1145) * st;
    }
    part[threadIdx.x] = local;
    __syncthreads();
    // inclusive scan of the per-thread partial sums (Hillis-Steele in LDS)
    for (int off = 1; off < kEnvBlock; off <<= 1) {
        double v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0.0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    double total = part[kEnvBlock - 1];
    double mean = total / (double)w;
    double run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0.0;
    for (int i = i0; i < i1; ++i) {
        float4 t = tex[(size_t)row * w + i];
        double lum = ((double)t.x * 0.2989 + (double)t.y * 0.5866 + (double)t.z * 0.1145) * st;
        run += lum;
        size_t k = (size_t)row * w + i;
        if (mean > 0.0) {
            cond_pdf[k] = (float)(lum / mean);
            cond_cdf[k] = i == w - 1 ? 1.0f : (float)(run / (double)w / mean);
        } else {
            cond_pdf[k] = 1.0f;
            cond_cdf[k] = i == w - 1 ? 1.0f : (float)((double)(i + 1) / (double)w);
        }
    }
    if (threadIdx.x == 0) row_mean[row] = mean;
}

// Marginal pdf/cdf over rows (CU/env_cameras.cu:59-74), one workgroup.
__global__ __launch_bounds__(kEnvBlock) void k_env_marginal(int h, const double* __restrict__ row_mean,
                                                            float* __restrict__ marg_pdf,
                                                            float* __restrict__ marg_cdf)
{
    __shared__ double part[kEnvBlock];
    int per = (h + kEnvBlock - 1) / kEnvBlock;
    int i0 = threadIdx.x * per, i1 = min(h, i0 + per);
    double local = 0.0;
    for (int i = i0; i < i1; ++i) local += row_mean[i];
    part[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < kEnvBlock; off <<= 1) {
        double v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0.0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    double mean = part[kEnvBlock - 1] / (double)h;
    double run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0.0;
    for (int i = i0; i < i1; ++i) {
        run += row_mean[i];
        marg_pdf[i] = mean > 0.0 ? (float)(row_mean[i] / mean) : 1.0f;
        marg_cdf[i] = i == h - 1 ? 1.0f : (mean > 0.0 ? (float)(run / (double)h / mean) : (float)((i + 1.0) / h));
    }
}

void build_env_tables(const float4* tex, int w, int h, float* marg_pdf, float* marg_cdf, float* cond_pdf,
                      float* cond_cdf, double* scratch, hipStream_t stream)
{
    hipLaunchKernelGGL(k_env_rows, dim3(h), dim3(kEnvBlock), 0, stream, tex, w, h, cond_pdf, cond_cdf, scratch);
    ORF_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(kEnvBlock), 0, stream, h, scratch, marg_pdf, marg_cdf);
    ORF_HIP_CHECK(hipGetLastError());
}

}  // namespace orf
