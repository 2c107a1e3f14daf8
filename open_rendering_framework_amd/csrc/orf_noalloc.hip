// orf_noalloc.hip -- orf_launch_noalloc: the megakernel launch of orf_launch
// for callers that cannot allocate device memory (code under graph capture,
// serving loops, harnesses that forbid hipMalloc).
//
// The caller passes the film (FP64 sums) and, optionally, the statistics
// counters; the only other device memory the launch needs -- the 8 per-XCD
// work-queue counters and, without caller counters, 4 statistics counters --
// lives in the file-scope array `scratch` below (128 bytes per device, held
// until the process ends).  Its device address is looked up once per device
// with hipGetSymbolAddress and used from then on; the host-side name of the
// array is never passed to a kernel.  Every call on a device shares that
// device's array and leaves its contents behind, so the counters are cleared
// by a kernel at the start of each call (no memset node: the call may be
// captured), and calls that overlap on one device race.
#include <mutex>

#include "orf_common.h"
#include "orf_context.h"
#include "orf_render.h"

namespace orf {

constexpr int kScratchWords = 16;  // [0..7] queue counters, [8..11] statistics fallback
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

namespace {

std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[64] = {nullptr};

// Device address of `scratch` on the current device (looked up on first use).
unsigned long long* scratch_on_current_device()
{
    int dev = 0;
    ORF_HIP_CHECK(hipGetDevice(&dev));
    ORF_REQUIRE(dev >= 0 && dev < 64, "orf_launch_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        ORF_REQUIRE(e == hipSuccess && p != nullptr,
                    std::string("orf_launch_noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned long long*)p;
    }
    return g_scratch_ptr[dev];
}

__global__ void k_clear_scratch(unsigned long long* p, int n)
{
    const int i = (int)threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

}  // namespace orf

extern "C" orf_status orf_launch_noalloc(orf_context ctx, int width, int height, uint32_t frame0, uint32_t spp,
                                         double* film, uint64_t* counters, void* stream)
{
    ORF_API_BEGIN
    orf_context_t* c = orf::ctx_check(ctx);
    ORF_REQUIRE(film != nullptr, "orf_launch_noalloc: null film");
    ORF_REQUIRE(width > 0 && height > 0 && width <= 16384 && height <= 16384, "orf_launch_noalloc: bad film size");
    ORF_REQUIRE(c->max_depth <= 4096, "orf_launch_noalloc: max_depth too large");
    ORF_REQUIRE(c->integrator == ORF_INTEGRATOR_PATH_TRACER,
                "orf_launch_noalloc: only the path tracer integrator runs without allocation");
    ORF_REQUIRE(orf::scene_ready(c),
                "orf_launch_noalloc: the scene has changes not yet on the device; call orf_accel_build first");
    ORF_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    unsigned long long* scr = orf::scratch_on_current_device();
    hipLaunchKernelGGL(orf::k_clear_scratch, dim3(1), dim3(64), 0, st, scr, orf::kScratchWords);
    ORF_HIP_CHECK(hipGetLastError());
    orf::CameraGPU cam = c->cam;
    if (!c->patch_set) {
        cam.patch_origin[0] = cam.patch_origin[1] = 0;
        cam.patch_dims[0] = (unsigned)width;
        cam.patch_dims[1] = (unsigned)height;
    }
    unsigned long long* stats = counters ? (unsigned long long*)counters : scr + 8;
    orf::FilmGPU f{film, stats, width, height, scr};
    orf::run_probe_pass(c, frame0, st);
    orf::launch_megakernel(orf::scene_view(c), cam, f, nullptr, frame0, spp, st);
    ORF_API_END
}
