// orf_bsdf.hip -- rough-conductor BSDF operator over SoA query batches (gfx950).
//
// One lane per query: 24 B in (wo xyz, alpha, u1, u2) and 28 B out (wi xyz,
// pdf, RGB weight), all unit-stride FP32 streams, so a wave64 issues fully
// coalesced 256 B loads/stores per stream.  The kernel is HBM-bound when the
// per-query arithmetic (two acos/sincos/atan2 chains, Smith masking twice, one
// FP32 complex Fresnel per distinct channel) fits under the memory time.
#include <cmath>
#include <cstring>

#include "orf_common.h"
#include "orf_fresnel.h"
#include "orf_microfacet.h"
#include "orf_render.h"
#include "orf_shade.h"

namespace orf {

__global__ __launch_bounds__(256) void k_ggx_conductor(int64_t n, const float* __restrict__ wx,
                                                       const float* __restrict__ wy,
                                                       const float* __restrict__ wz,
                                                       const float* __restrict__ alpha,
                                                       const float* __restrict__ u1,
                                                       const float* __restrict__ u2, ConductorEta eta,
                                                       float* __restrict__ owx, float* __restrict__ owy,
                                                       float* __restrict__ owz, float* __restrict__ opdf,
                                                       float* __restrict__ ow_r, float* __restrict__ ow_g,
                                                       float* __restrict__ ow_b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 wo = mk3(wx[i], wy[i], wz[i]);
    float a = alpha[i];
    const f3 nrm = mk3(0.0f, 0.0f, 1.0f);
    f3 ffn = faceforward(nrm, wo, nrm);
    f3 wi, w;
    float pdf;
    ggx_conductor_sample(wo, ffn, a, a, u1[i], u2[i], eta, wi, pdf, w);
    owx[i] = wi.x;
    owy[i] = wi.y;
    owz[i] = wi.z;
    opdf[i] = pdf;
    ow_r[i] = w.x;
    ow_g[i] = w.y;
    ow_b[i] = w.z;
}

}  // namespace orf

extern "C" orf_status orf_bsdf_ggx_conductor_batch(int64_t n, const float* wo_x, const float* wo_y,
                                                   const float* wo_z, const float* alpha,
                                                   const float* u1, const float* u2,
                                                   const double ior_re[3], const double ior_im[3],
                                                   float* wi_x, float* wi_y, float* wi_z, float* pdf,
                                                   float* w_r, float* w_g, float* w_b, void* stream)
{
    ORF_API_BEGIN
    ORF_REQUIRE(n >= 0, "orf_bsdf_ggx_conductor_batch: n < 0");
    if (n == 0) return ORF_SUCCESS;
    ORF_REQUIRE(wo_x && wo_y && wo_z && alpha && u1 && u2 && wi_x && wi_y && wi_z && pdf && w_r &&
                    w_g && w_b,
                "orf_bsdf_ggx_conductor_batch: null array");
    ORF_REQUIRE(ior_re && ior_im, "orf_bsdf_ggx_conductor_batch: null ior");
    orf::ConductorEta eta;
    orf::conductor_eta_host(ior_re, ior_im, eta);
    const int block = 256;
    int64_t grid = (n + block - 1) / block;
    ORF_REQUIRE(grid <= 0x7fffffffLL, "orf_bsdf_ggx_conductor_batch: batch too large");
    hipLaunchKernelGGL(orf::k_ggx_conductor, dim3((unsigned)grid), dim3(block), 0,
                       (hipStream_t)stream, n, wo_x, wo_y, wo_z, alpha, u1, u2, eta, wi_x, wi_y,
                       wi_z, pdf, w_r, w_g, w_b);
    ORF_HIP_CHECK(hipGetLastError());
    ORF_API_END
}
