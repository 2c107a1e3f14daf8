/*
 * orf_oracle.h -- CPU restatement of the reference render path (TEST INFRASTRUCTURE).
 *
 * This directory is the oracle: a plain-C restatement of the algorithms of
 * andrea-luongo/open_rendering_framework's render hot path, written from reading
 * the reference sources (each function cites the file:line it follows; paths are
 * relative to the reference root, ORF/ = open_rendering_framework/,
 * CU/ = open_rendering_framework/CUDA_files/).  Only tests/, the smoke() check
 * and bench.py's cpu_baseline leg may load it, and only as the checker or the
 * CPU baseline.  The product (open_rendering_framework_amd) never links it.
 *
 * Precision: the scalar type R is float by default (the reference's device math
 * is FP32 with an FP64 complex Fresnel), or double when built with
 * -DORACLE_FP64 (the FP64 oracle used for converged-image RMSE).  <tgmath.h>
 * dispatches every libm call to the float or double routine by argument type.
 *
 * Pinning (what makes this oracle trustworthy) lives in tests/test_oracle_*.py:
 * published PCG32/TEA/LCG vectors, closed-form microfacet integrals, textbook
 * Fresnel special cases, chi-square tests of the samplers, white-furnace and
 * brute-force BVH checks.  Functions without such a check are marked
 * "parity unpinned" in their comment.
 */
#ifndef ORF_ORACLE_H
#define ORF_ORACLE_H

#include <stdint.h>

#ifdef ORACLE_FP64
typedef double R;
#else
typedef float R;
#endif

typedef struct { R x, y; } v2;
typedef struct { R x, y, z; } v3;

/* MyComplex / MyComplex3 (ORF/MyComplex.h:7-16): parts are double. */
typedef struct { double re, im; } cplx;
typedef struct { cplx x, y, z; } cplx3;

/* ---- random numbers (ORF/random.h) ---- */
uint32_t orc_tea_key(uint32_t v0, uint32_t v1, uint32_t rounds, const uint32_t key[4]);
uint32_t orc_tea16(uint32_t v0, uint32_t v1);
R orc_rnd_tea(uint32_t *prev);
uint32_t orc_lcg(uint32_t *prev);
R orc_rnd(uint32_t *prev);
/* MD5 (ORF/md5.h) of a message of at most 55 bytes; rnd_accurate (ORF/random.h:134-142)
 * on a two-word state. */
void orc_md5(const uint8_t* msg, uint32_t len, uint32_t out[4]);
R orc_rnd_accurate(uint32_t s[2]);
/* A path's random state: the TEA state every program draws from (first
 * member: the shaders receive a pointer to it) and the 64-bit state of the
 * translucent programs (PerRayData_radiance seed / seed64, ORF/structs.h:20-28). */
typedef struct {
    uint32_t tea;
    uint32_t s64[2];
} orc_rng;
/* PCG32 (O'Neill 2014, pcg32_srandom_r / pcg32_random_r), used by the
 * synthetic BSDF-query generator; not in the reference. */
typedef struct { uint64_t state, inc; } orc_pcg32;
void orc_pcg32_seed(orc_pcg32 *r, uint64_t initstate, uint64_t initseq);
uint32_t orc_pcg32_next(orc_pcg32 *r);

/* ---- vector helpers ---- */
static inline v3 V3(R x, R y, R z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, R s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline R vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b)
{
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* ---- microfacet / sampling / Fresnel (oracle_math.c) ---- */
v3 orc_normalize(v3 v);
v3 orc_faceforward(v3 n, v3 i, v3 nref);
void orc_create_onb(v3 n, v3 *U, v3 *V);
v3 orc_to_local(v3 w, v3 n);
v3 orc_to_world(v3 w, v3 n);
void orc_rotate_to_normal(v3 n, v3 *v);
v3 orc_sample_cosine_weighted(v3 n, uint32_t *t);
v3 orc_sample_barycentric(uint32_t *t);

R orc_ggx_alpha_i(v3 wi, R ax, R ay);
R orc_ggx_G1_cos(R cos_theta, R alpha_sqr);
R orc_ggx_G1(v3 v, v3 m, v3 n, R ax, R ay);
R orc_ggx_projected_area(v3 wi, R ax, R ay);
R orc_ggx_eval_ndf(v3 wm, R ax, R ay);
R orc_ggx_eval_vndf(v3 wi, v3 wm, R ax, R ay);
v2 orc_ggx_sample_p22_11(R theta_i, R z1, R z2);
v3 orc_ggx_sample_vndf(v3 wi, R ax, R ay, R z1, R z2);

cplx orc_cdiv_fc(R c1, cplx c2);
v3 orc_fresnel_conductor(v3 w_i, v3 w_m, const cplx3 *eta);
R orc_fresnel_dielectric(R cos_theta, R ior1_over_ior2);

double orc_erfinv(double x);
R orc_beckmann_alpha_i(v3 wi, R ax, R ay);
R orc_beckmann_G1_cos(R cos_theta, R alpha_sqr);
R orc_beckmann_G1(v3 v, v3 m, v3 n, R ax, R ay);
R orc_beckmann_projected_area(v3 wi, R ax, R ay);
R orc_beckmann_eval_ndf(v3 wm, R ax, R ay);
R orc_beckmann_eval_vndf(v3 wi, v3 wm, R ax, R ay);
v2 orc_beckmann_sample_p22_11(R theta_i, R z1, R z2);
v3 orc_beckmann_sample_vndf(v3 wi, R ax, R ay, R z1, R z2);
R orc_fresnel_R3(R cos1, R cos2, R eta);
/* microsurface multiple scattering (Heitz et al. 2016) */
R orc_ggx_smith_lambda(v3 wi, R ax, R ay);
R orc_height_cdf(R h);
R orc_height_inv_cdf(R z);
R orc_ggx_masking_height(v3 wi, R h, R ax, R ay);
R orc_ggx_sample_height(v3 wr, R hr, R z, R ax, R ay);
int orc_ggx_ms_dielectric_sample(v3 local_wi, R eta, R ax, R ay, uint32_t* seed, int cap, v3* w_out, int* order);
int orc_beckmann_ms_dielectric_sample(v3 local_wi, R eta, R ax, R ay, uint32_t* seed, int cap, v3* w_out, int* order);
int orc_beckmann_ms_conductor_sample(v3 local_wi, const cplx3* eta, const cplx3* ieta, R ax, R ay, uint32_t* seed,
                                     int cap, v3* w_out, v3* weight, int* order);
int orc_ggx_ms_conductor_sample(v3 local_wi, const cplx3* eta, const cplx3* ieta, R ax, R ay, uint32_t* seed, int cap,
                                v3* w_out, v3* weight, int* order);
R orc_fresnel_R_dir(v3 w_i, v3 w_m, R eta);
R orc_ggx_eval_dielectric_phase(v3 w_i, v3 w_o, int wi_out, int wo_out, R m_eta, R ax, R ay);
R orc_ggx_ms_dielectric_eval(v3 w_i, v3 w_o, R eta, R ax, R ay, uint32_t* seed, int cap);
R orc_ms_masking_height(int dist, v3 wi, R h, R ax, R ay);
void orc_ms_diffuse_sample(int dist, v3 local_wi, R ax, R ay, v3 rho, uint32_t* seed, int cap, v3* w_out, v3* w_m,
                           v3* weight, int* order);
v3 orc_ms_diffuse_eval(int dist, v3 w_i, v3 w_o, R ax, R ay, v3 rho, uint32_t* seed, int cap);
v3 orc_ms_conductor_eval(int dist, v3 w_i, v3 w_o, const cplx3* eta, const cplx3* ieta, R ax, R ay, uint32_t* seed,
                         int cap);
R orc_beckmann_eval_dielectric_phase(v3 w_i, v3 w_o, int wi_out, int wo_out, R m_eta, R ax, R ay);
R orc_beckmann_ms_dielectric_eval(v3 w_i, v3 w_o, R eta, R ax, R ay, uint32_t* seed, int cap);
/* Walter et al. 2007 model and the visible-normal density (ORF/Microfacet.h). */
R orc_masking_iso(v3 v, v3 m, v3 n, R a, int dist);
v3 orc_sample_normal_walter(v3 n, R z1, R z2, R a, int dist);
R orc_eval_visible_normal(v3 wi, v3 wm, v3 n, R ax, R ay, int dist);
void orc_ggx_ms_eval_batch(int n, const float* wi, const float* wo, float eta, float alpha, uint32_t seed, int cap,
                           float* out);
void orc_beckmann_ms_eval_batch(int n, const float* wi, const float* wo, float eta, float alpha, uint32_t seed, int cap,
                                float* out);
void orc_ms_diffuse_sample_batch(int n, const float* wi, float alpha, int dist, const float* rho, uint32_t seed, int cap,
                                 float* wo, float* wm, float* weight, int32_t* order);
void orc_ms_diffuse_eval_batch(int n, const float* wi, const float* wo, float alpha, int dist, const float* rho,
                               uint32_t seed, int cap, float* out);
void orc_ms_conductor_eval_batch(int n, const float* wi, const float* wo, const double* ior, float alpha, int dist,
                                 uint32_t seed, int cap, float* out);
void orc_walter_sample_batch(int n, const float* z1, const float* z2, float alpha, int dist, float* m);
void orc_reflect_weight_batch(int n, const float* wi, const float* z1, const float* z2, float alpha, int dist,
                              int model, float* out);

/* One query of the rough-conductor operator (CU/metallic_shader.cu:84-158,
 * visible-normal branch, GGX, surface normal (0,0,1)).  Returns 1 when the
 * path continues, 0 when the reference would stop (G1 <= 0). */
int orc_ggx_conductor_query(v3 wo, R alpha, R u1, R u2, const cplx3 *eta,
                            v3 *wi, R *pdf, v3 *weight);

#endif
