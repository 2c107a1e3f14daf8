// orf_depth.hip -- depth integrator (gfx950): the DepthTracer ray generation
// program (CU/depth_tracer.cu:33-58) with the depth closest hit
// (CU/depth_shader.cu:13-19).  Per sample: the face-forwarded shading normal
// mapped as (1 + n) / 2 and the hit distance (both 0 where the ray leaves the
// scene or meets a disk/sphere emitter, whose materials have no depth
// program), accumulated into the FP64 film (RGB sums, finite count) and a FP64
// depth sum per pixel.  One wave64 workgroup per 8x8 pixel tile, one lane per
// pixel, every sample of the launch in that lane.
#include "orf_common.h"
#include "orf_path.h"
#include "orf_render.h"

namespace orf {

constexpr int kDepthBlock = 64;

__global__ __launch_bounds__(kDepthBlock) void k_depth(SceneGPU sc, CameraGPU cam, FilmGPU film,
                                                       double* __restrict__ depth, unsigned frame0, unsigned spp)
{
    __shared__ unsigned stack_lds[kLdsStack * kDepthBlock];
    unsigned* stk = stack_lds + threadIdx.x;
    const int w = film.width, h = film.height;
    const int tiles_x = (w + 7) / 8;
    const int x = (int)(blockIdx.x % (unsigned)tiles_x) * 8 + (int)(threadIdx.x & 7);
    const int y = (int)(blockIdx.x / (unsigned)tiles_x) * 8 + (int)(threadIdx.x >> 3);
    if (x >= w || y >= h) return;
    const size_t px = (size_t)y * w + x;
    double r = 0.0, g = 0.0, b = 0.0, dsum = 0.0;
    unsigned cnt = 0, bad = 0;
    for (unsigned s = 0; s < spp; ++s) {
        f3 dir;
        uint32_t rng = camera_ray(cam, w, h, x, y, frame0 + s, dir);
        PathState ps;
        path_begin(ps, cam.eye, dir, rng, seed64(0u));
        Hit hit;
        f3 n = mk3(0.0f);
        float t = 0.0f;
        if (trace<false>(sc, cam.eye, dir, sc.scene_epsilon, kRayMax, hit, stk, kDepthBlock) && hit.leaf >= 0) {
            n = hit_geometry(sc, ps, hit).ffn;
            t = hit.t;
        }
        if (!isfinite(t)) {
            bad++;
            continue;
        }
        r += (double)((1.0f + n.x) * 0.5f);
        g += (double)((1.0f + n.y) * 0.5f);
        b += (double)((1.0f + n.z) * 0.5f);
        dsum += (double)t;
        cnt++;
    }
    film.sums[4 * px + 0] += r;
    film.sums[4 * px + 1] += g;
    film.sums[4 * px + 2] += b;
    film.sums[4 * px + 3] += (double)cnt;
    depth[px] += dsum;
    atomicAdd(film.counters + 1, (unsigned long long)spp);
    atomicAdd(film.counters + 3, (unsigned long long)spp);
    if (bad) atomicAdd(film.counters, (unsigned long long)bad);
}

void launch_depth(const SceneGPU& sc, const CameraGPU& cam, const FilmGPU& film, double* depth, unsigned frame0,
                  unsigned spp, hipStream_t stream)
{
    const long long tiles = (long long)((film.width + 7) / 8) * ((film.height + 7) / 8);
    ORF_REQUIRE(tiles > 0 && tiles < 0x7fffffffLL && depth, "depth launch: bad film");
    hipLaunchKernelGGL(k_depth, dim3((unsigned)tiles), dim3(kDepthBlock), 0, stream, sc, cam, film, depth, frame0,
                       spp);
    ORF_HIP_CHECK(hipGetLastError());
}

}  // namespace orf
