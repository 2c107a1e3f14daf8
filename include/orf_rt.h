/*
 * orf_rt.h -- C-ABI of the MI355X-native render runtime (liborf_rt.so).
 *
 * This is the drop-in boundary that replaces the OptiX 5 host runtime the
 * reference's scene classes call (optixu/optixpp_namespace.h as used from
 * ORF/OptixScene.cpp, ORF/OptixSceneLoader.cpp, ORF/...Material.cpp,
 * ORF/Light.cpp, ORF/Camera.cpp, ORF/Integrator.cpp, ORF/Background.cpp,
 * ORF/ObjLoader.cpp).  Plain C: opaque handles, plain pointers and sizes, one
 * status code per call and orf_last_error() for the message (the role of
 * RTresult + optix::Exception).  Device pointers are HIP device addresses;
 * `stream` is a hipStream_t (NULL = the default stream).
 *
 * Threading: one host thread owns a context (as one Qt thread owned the
 * optix::Context); calls on different contexts may run concurrently.
 */
#ifndef ORF_RT_H
#define ORF_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum orf_status {
    ORF_SUCCESS = 0,
    ORF_ERROR_INVALID_VALUE = 1,  /* bad argument / handle / variable name  */
    ORF_ERROR_HIP = 2,            /* HIP runtime failure                     */
    ORF_ERROR_OUT_OF_MEMORY = 3,  /* host or device allocation failed        */
    ORF_ERROR_NOT_FOUND = 4,      /* file / program / variable not found     */
    ORF_ERROR_FILE_FORMAT = 5,    /* malformed OBJ / HDR / JSON / raw film   */
    ORF_ERROR_UNKNOWN = 99
} orf_status;

/* Text of the last error raised on the calling thread ("" if none). */
const char* orf_last_error(void);
/* Library version string, e.g. "orf_rt 0.1 gfx950". */
const char* orf_version(void);

/* ------------------------------------------------------------------------
 * Rough-conductor BSDF operator (batch of independent queries).
 *
 * Replaces the visible-normal sampling step of the Metallic material's
 * closest-hit program, CU/metallic_shader.cu:84-158 (microfacet normal from
 * ORF/Microfacet.h:72-83 -> ORF/MicrofacetGGX.h:190-207, masking
 * ORF/MicrofacetGGX.h:63-74, Fresnel ORF/Fresnel.h:78-91), with the material
 * parameters MetallicMaterial sets (ior, roughness: ORF/MetallicMaterial.cpp).
 * There, one query ran per radiance hit inside OptiX traversal; here it is a
 * flat SoA batch in local shading space (macro-normal +z).
 *
 * Inputs (device, FP32, n each): wo (towards the viewer, unit, z > 0),
 * isotropic GGX roughness alpha, uniform numbers u1 u2.  ior_re/ior_im
 * (host, 3 each): complex IOR per RGB channel, relative to air (n1 = 1).
 * Outputs (device, FP32, n each): wi = sampled reflection direction,
 * pdf = D_wo(m) / (4 |wi.m|) in solid angle, weight = F(wo.m) G1(wi)
 * (= f cos / pdf).  weight is 0 where the reference stops the path.
 * ------------------------------------------------------------------------ */
orf_status orf_bsdf_ggx_conductor_batch(int64_t n,
                                        const float* wo_x, const float* wo_y, const float* wo_z,
                                        const float* alpha, const float* u1, const float* u2,
                                        const double ior_re[3], const double ior_im[3],
                                        float* wi_x, float* wi_y, float* wi_z, float* pdf,
                                        float* w_r, float* w_g, float* w_b,
                                        void* stream);

/* Host-side generator of the synthetic operator workload: one PCG32 stream
 * (O'Neill 2014, initstate seed, stream seq), five draws per query:
 * cos(theta)=u0, phi=2 pi u1, alpha=0.05+0.95 u2, u1=u3, u2=u4.  Host arrays. */
orf_status orf_gen_bsdf_queries(int64_t n, uint64_t seed, uint64_t seq,
                                float* wo_x, float* wo_y, float* wo_z, float* alpha,
                                float* u1, float* u2);

#ifdef __cplusplus
}
#endif

#endif /* ORF_RT_H */
