// orf_render.hip -- plain megakernel path tracer and environment CDF build (gfx950).
//
// Megakernel: one wave64 workgroup per (8x8 pixel tile, sample chunk); each lane
// walks all paths of its pixel's sample range (path loop in orf_path.h) and adds
// FP64 sums of its finite samples to the film (plain RMW when one chunk covers
// all samples, else FP64 atomics).
#include <algorithm>

#include "orf_common.h"
#include "orf_path.h"
#include "orf_render.h"

namespace orf {

constexpr int kMegaBlock = 64;

__global__ __launch_bounds__(kMegaBlock, ORF_MEGA_WAVES) void k_render_mega(SceneGPU sc, CameraGPU cam, FilmGPU film,
                                                            const int* remap, unsigned frame0, unsigned spp,
                                                            int chunks, unsigned per, int tiles_x)
{
    __shared__ unsigned stack_lds[kLdsStack * kMegaBlock];
    unsigned* stk = stack_lds + threadIdx.x;
    unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    int tile = blk / chunks, chunk = blk % chunks;
    int x = (tile % tiles_x) * 8 + (threadIdx.x & 7);
    int y = (tile / tiles_x) * 8 + (threadIdx.x >> 3);
    if (x >= film.width || y >= film.height) return;
    // local sample range of this chunk (frame = frame0 + k wraps like the seeds do)
    unsigned k0 = (unsigned)chunk * per;
    unsigned k1 = min(k0 + per, spp);
    if (k0 >= k1) return;
    bool inside = !((unsigned)x < cam.patch_origin[0] || (unsigned)x > cam.patch_origin[0] + cam.patch_dims[0] ||
                    (unsigned)y < cam.patch_origin[1] || (unsigned)y > cam.patch_origin[1] + cam.patch_dims[1]);
    double sr = 0.0, sg = 0.0, sb = 0.0;
    unsigned cnt = 0, bad = 0;
    PathStats st;
    for (unsigned k = k0; k < k1; ++k) {
        f3 L = mk3(0.0f);
        if (inside) {
            f3 dir;
            uint32_t rng = camera_ray(cam, film.width, film.height, x, y, frame0 + k, dir);
            L = path_radiance(sc, remap, cam.eye, dir, rng, stk, kMegaBlock, st);
        }
        if (finite3(L)) {
            sr += (double)L.x;
            sg += (double)L.y;
            sb += (double)L.z;
            cnt++;
        } else {
            bad++;
        }
    }
    double* px = film.sums + 4 * ((size_t)y * film.width + x);
    if (chunks == 1) {
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
    if (bad) atomicAdd(film.counters + 0, (unsigned long long)bad);
    atomicAdd(film.counters + 1, (unsigned long long)st.ext);
    atomicAdd(film.counters + 2, (unsigned long long)st.shd);
    atomicAdd(film.counters + 3, (unsigned long long)(k1 - k0));
}

void launch_megakernel(const SceneGPU& sc, const CameraGPU& cam, const FilmGPU& film, const int* remap,
                       unsigned frame0, unsigned spp, hipStream_t stream)
{
    if (spp == 0 || film.width <= 0 || film.height <= 0) return;
    int tiles_x = (film.width + 7) / 8, tiles_y = (film.height + 7) / 8;
    int64_t tiles = (int64_t)tiles_x * tiles_y;
    // enough lanes to fill 256 CUs several times over; each lane runs >= 4 samples
    int64_t target_waves = 256 * 64;
    int chunks = (int)std::max<int64_t>(1, std::min<int64_t>((target_waves + tiles - 1) / tiles, (spp + 3) / 4));
    unsigned per = (spp + chunks - 1) / chunks;
    chunks = (int)((spp + per - 1) / per);
    int64_t blocks = tiles * chunks;
    ORF_REQUIRE(blocks < 0x7fffffffLL, "launch: image too large");
    hipLaunchKernelGGL(k_render_mega, dim3((unsigned)blocks), dim3(kMegaBlock), 0, stream, sc, cam, film, remap,
                       frame0, spp, chunks, per, tiles_x);
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
