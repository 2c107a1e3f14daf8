/*
 * oracle_batch.c -- array entry points of the oracle for ctypes (TEST INFRASTRUCTURE).
 *
 * Thin loops over the scalar restatements in oracle_math.c so that tests can
 * compare whole batches with the product, and so the benchmark can time the
 * CPU restatement of the operator (OpenMP over queries, static schedule).
 */
#include <stdint.h>
#include <stdlib.h>
#include <tgmath.h>
#include "orf_oracle.h"

void orc_tea_stream(uint32_t seed, int n, uint32_t *out_state, float *out_u)
{
    uint32_t s = seed;
    for (int i = 0; i < n; i++) {
        R u = orc_rnd_tea(&s);
        out_state[i] = s;
        out_u[i] = (float)u;
    }
}

/* n draws of rnd_accurate from the state (x, y): states after each draw (2 words) and values. */
void orc_accurate_stream(uint32_t x, uint32_t y, int n, uint32_t *out_state, float *out_u)
{
    uint32_t st[2] = {x, y};
    for (int i = 0; i < n; i++) {
        R u = orc_rnd_accurate(st);
        out_state[2 * i] = st[0];
        out_state[2 * i + 1] = st[1];
        out_u[i] = (float)u;
    }
}

void orc_tea_batch(const uint32_t *v0, const uint32_t *v1, int n, uint32_t rounds,
                   const uint32_t *key, uint32_t *out)
{
    for (int i = 0; i < n; i++)
        out[i] = orc_tea_key(v0[i], v1[i], rounds, key);
}

void orc_lcg_stream(uint32_t seed, int n, uint32_t *out_state, float *out_u)
{
    uint32_t s = seed;
    for (int i = 0; i < n; i++) {
        R u = orc_rnd(&s);
        out_state[i] = s;
        out_u[i] = (float)u;
    }
}

void orc_pcg32_stream(uint64_t initstate, uint64_t initseq, int n, uint32_t *out)
{
    orc_pcg32 r;
    orc_pcg32_seed(&r, initstate, initseq);
    for (int i = 0; i < n; i++)
        out[i] = orc_pcg32_next(&r);
}

/* Synthetic operator queries: one PCG32 stream, five draws per query,
 * u = (x >> 8) * 2^-24.  cos(theta) = u0, phi = 2 pi u1 (uniform by solid
 * angle over the upper hemisphere), alpha = 0.05 + 0.95 u2, BSDF numbers
 * u3, u4.  Arithmetic is float with float libm. */
void orc_gen_bsdf_queries(int64_t n, uint64_t seed, uint64_t seq,
                          float *wx, float *wy, float *wz, float *alpha, float *u1, float *u2)
{
    orc_pcg32 r;
    orc_pcg32_seed(&r, seed, seq);
    const float two_pi = 6.28318530717958647692f;
    for (int64_t i = 0; i < n; i++) {
        float u[5];
        for (int k = 0; k < 5; k++)
            u[k] = (float)(orc_pcg32_next(&r) >> 8) * (1.0f / 16777216.0f);
        float c = u[0];
        float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
        float phi = two_pi * u[1];
        wx[i] = s * cosf(phi);
        wy[i] = s * sinf(phi);
        wz[i] = c;
        alpha[i] = 0.05f + 0.95f * u[2];
        u1[i] = u[3];
        u2[i] = u[4];
    }
}

static cplx3 eta_from_ior(const double *ior_re, const double *ior_im)
{
    cplx3 e;
    cplx i0 = {ior_re[0], ior_im[0]}, i1 = {ior_re[1], ior_im[1]}, i2 = {ior_re[2], ior_im[2]};
    e.x = orc_cdiv_fc((R)1.0f, i0);
    e.y = orc_cdiv_fc((R)1.0f, i1);
    e.z = orc_cdiv_fc((R)1.0f, i2);
    return e;
}

/* eta = n1 / ior with n1 = 1 (CU/metallic_shader.cu:59-61). */
void orc_conductor_eta(const double *ior_re, const double *ior_im, double *eta_re, double *eta_im)
{
    cplx3 e = eta_from_ior(ior_re, ior_im);
    eta_re[0] = e.x.re; eta_im[0] = e.x.im;
    eta_re[1] = e.y.re; eta_im[1] = e.y.im;
    eta_re[2] = e.z.re; eta_im[2] = e.z.im;
}

/* The rough-conductor operator over a batch (SoA in, SoA out). */
void orc_bsdf_ggx_conductor_batch(int64_t n, const float *wx, const float *wy, const float *wz,
                                  const float *alpha, const float *u1, const float *u2,
                                  const double *ior_re, const double *ior_im,
                                  float *owx, float *owy, float *owz, float *opdf,
                                  float *or_, float *og, float *ob, int nthreads)
{
    cplx3 eta = eta_from_ior(ior_re, ior_im);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; i++) {
        v3 wi, w;
        R pdf;
        orc_ggx_conductor_query(V3(wx[i], wy[i], wz[i]), alpha[i], u1[i], u2[i], &eta, &wi, &pdf, &w);
        owx[i] = (float)wi.x; owy[i] = (float)wi.y; owz[i] = (float)wi.z;
        opdf[i] = (float)pdf;
        or_[i] = (float)w.x; og[i] = (float)w.y; ob[i] = (float)w.z;
    }
}

/* Vectorised single functions for the closed-form pinning tests. */
void orc_ggx_G1_batch(int n, const float *v, const float *m, const float *nrm, const float *ax,
                      const float *ay, float *out)
{
    for (int i = 0; i < n; i++)
        out[i] = (float)orc_ggx_G1(V3(v[3 * i], v[3 * i + 1], v[3 * i + 2]),
                                   V3(m[3 * i], m[3 * i + 1], m[3 * i + 2]),
                                   V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), ax[i], ay[i]);
}

void orc_ggx_ndf_batch(int n, const double *wm, double ax, double ay, double *out)
{
    for (int i = 0; i < n; i++)
        out[i] = (double)orc_ggx_eval_ndf(V3((R)wm[3 * i], (R)wm[3 * i + 1], (R)wm[3 * i + 2]), (R)ax, (R)ay);
}

void orc_ggx_vndf_batch(int n, const double *wi, const double *wm, double ax, double ay, double *out)
{
    for (int i = 0; i < n; i++)
        out[i] = (double)orc_ggx_eval_vndf(V3((R)wi[0], (R)wi[1], (R)wi[2]),
                                           V3((R)wm[3 * i], (R)wm[3 * i + 1], (R)wm[3 * i + 2]), (R)ax, (R)ay);
}

void orc_ggx_sample_vndf_batch(int n, const float *wi, float ax, float ay, const float *z1,
                               const float *z2, float *out)
{
    for (int i = 0; i < n; i++) {
        v3 m = orc_ggx_sample_vndf(V3(wi[0], wi[1], wi[2]), ax, ay, z1[i], z2[i]);
        out[3 * i] = (float)m.x; out[3 * i + 1] = (float)m.y; out[3 * i + 2] = (float)m.z;
    }
}

void orc_fresnel_conductor_batch(int n, const float *cos_i, const double *ior_re,
                                 const double *ior_im, float *out)
{
    cplx3 eta = eta_from_ior(ior_re, ior_im);
    for (int i = 0; i < n; i++) {
        R c = cos_i[i];
        /* w_m = (0,0,1), w_i in the xz plane with the requested cosine */
        R s = sqrt((R)1 - c * c);
        v3 F = orc_fresnel_conductor(V3(s, 0, c), V3(0, 0, 1), &eta);
        out[3 * i] = (float)F.x; out[3 * i + 1] = (float)F.y; out[3 * i + 2] = (float)F.z;
    }
}

void orc_fresnel_dielectric_batch(int n, const float *cos_i, float eta, float *out)
{
    for (int i = 0; i < n; i++)
        out[i] = (float)orc_fresnel_dielectric(cos_i[i], eta);
}

void orc_cosine_batch(int n, const float *nrm, uint32_t seed, float *out)
{
    uint32_t t = seed;
    v3 nn = V3(nrm[0], nrm[1], nrm[2]);
    for (int i = 0; i < n; i++) {
        v3 d = orc_sample_cosine_weighted(nn, &t);
        out[3 * i] = (float)d.x; out[3 * i + 1] = (float)d.y; out[3 * i + 2] = (float)d.z;
    }
}

/* Beckmann restatements, vectorised for the closed-form pinning tests. */
void orc_beckmann_ndf_batch(int n, const double *wm, double ax, double ay, double *out)
{
    for (int i = 0; i < n; i++)
        out[i] = (double)orc_beckmann_eval_ndf(V3((R)wm[3 * i], (R)wm[3 * i + 1], (R)wm[3 * i + 2]), (R)ax, (R)ay);
}

void orc_beckmann_vndf_batch(int n, const double *wi, const double *wm, double ax, double ay, double *out)
{
    for (int i = 0; i < n; i++)
        out[i] = (double)orc_beckmann_eval_vndf(V3((R)wi[0], (R)wi[1], (R)wi[2]),
                                                V3((R)wm[3 * i], (R)wm[3 * i + 1], (R)wm[3 * i + 2]), (R)ax, (R)ay);
}

void orc_beckmann_proj_batch(int n, const double *wi, double ax, double ay, double *out)
{
    for (int i = 0; i < n; i++)
        out[i] = (double)orc_beckmann_projected_area(V3((R)wi[3 * i], (R)wi[3 * i + 1], (R)wi[3 * i + 2]), (R)ax,
                                                     (R)ay);
}

void orc_beckmann_G1_cos_batch(int n, const double *cos_t, double alpha, double *out)
{
    for (int i = 0; i < n; i++) out[i] = (double)orc_beckmann_G1_cos((R)cos_t[i], (R)(alpha * alpha));
}

void orc_beckmann_sample_vndf_batch(int n, const float *wi, float ax, float ay, const float *z1, const float *z2,
                                    float *out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        v3 m = orc_beckmann_sample_vndf(V3(wi[0], wi[1], wi[2]), ax, ay, z1[i], z2[i]);
        out[3 * i] = (float)m.x; out[3 * i + 1] = (float)m.y; out[3 * i + 2] = (float)m.z;
    }
}

void orc_erfinv_batch(int n, const double *x, double *out)
{
    for (int i = 0; i < n; i++) out[i] = orc_erfinv(x[i]);
}

/* Microsurface walks from one local direction: exit direction, event count and
 * completion flag (0: exceeded `cap`), one TEA stream per walk (tea16(i, seed)). */
void orc_ggx_ms_sample_batch(int n, const float* wi, float eta, float alpha, uint32_t seed, int cap, float* wo,
                             int32_t* order, int32_t* ok)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        uint32_t t = orc_tea16((uint32_t)i, seed);
        v3 out = V3(0, 0, 0);
        int o = 0;
        ok[i] = orc_ggx_ms_dielectric_sample(V3(wi[0], wi[1], wi[2]), eta, alpha, alpha, &t, cap, &out, &o);
        order[i] = o;
        wo[3 * i] = (float)out.x;
        wo[3 * i + 1] = (float)out.y;
        wo[3 * i + 2] = (float)out.z;
    }
}

/* Conductor microsurface walks (dist 1 GGX, 0 Beckmann) from the fixed local wi (stream tea16(i, seed)),
 * complex ior (re, im) per channel given as 6 doubles: exit direction, weight
 * (3 per walk), order and ok. */
void orc_ggx_ms_conductor_batch(int n, const float* wi, const double* ior, float alpha, int dist, uint32_t seed,
                                int cap, float* wo, float* weight, int32_t* order, int32_t* ok)
{
    cplx3 eta, ieta;
    ieta.x = (cplx){ior[0], ior[1]};
    ieta.y = (cplx){ior[2], ior[3]};
    ieta.z = (cplx){ior[4], ior[5]};
    eta.x = orc_cdiv_fc((R)1.0f, ieta.x);
    eta.y = orc_cdiv_fc((R)1.0f, ieta.y);
    eta.z = orc_cdiv_fc((R)1.0f, ieta.z);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        uint32_t t = orc_tea16((uint32_t)i, seed);
        v3 out = V3(0, 0, 0), w = V3(0, 0, 0);
        int o = 0;
        ok[i] = dist == 1 ? orc_ggx_ms_conductor_sample(V3(wi[0], wi[1], wi[2]), &eta, &ieta, alpha, alpha, &t, cap,
                                                        &out, &w, &o)
                          : orc_beckmann_ms_conductor_sample(V3(wi[0], wi[1], wi[2]), &eta, &ieta, alpha, alpha, &t,
                                                             cap, &out, &w, &o);
        order[i] = o;
        wo[3 * i] = (float)out.x;
        wo[3 * i + 1] = (float)out.y;
        wo[3 * i + 2] = (float)out.z;
        weight[3 * i] = (float)w.x;
        weight[3 * i + 1] = (float)w.y;
        weight[3 * i + 2] = (float)w.z;
    }
}

/* Beckmann walks from the fixed local wi (stream tea16(i, seed)): dielectric
 * (relative ior eta) exit directions, orders and flags. */
void orc_beckmann_ms_sample_batch(int n, const float* wi, float eta, float alpha, uint32_t seed, int cap, float* wo,
                                  int32_t* order, int32_t* ok)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        uint32_t t = orc_tea16((uint32_t)i, seed);
        v3 out = V3(0, 0, 0);
        int o = 0;
        ok[i] = orc_beckmann_ms_dielectric_sample(V3(wi[0], wi[1], wi[2]), eta, alpha, alpha, &t, cap, &out, &o);
        order[i] = o;
        wo[3 * i] = (float)out.x;
        wo[3 * i + 1] = (float)out.y;
        wo[3 * i + 2] = (float)out.z;
    }
}

/* One multiple-scattering BSDF x cos estimate per direction pair wo[i] for the
 * fixed local wi (stream tea16(i, seed)), at most `cap` events. */
void orc_ggx_ms_eval_batch(int n, const float* wi, const float* wo, float eta, float alpha, uint32_t seed, int cap,
                           float* out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        uint32_t t = orc_tea16((uint32_t)i, seed);
        out[i] = (float)orc_ggx_ms_dielectric_eval(V3(wi[0], wi[1], wi[2]), V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]),
                                                   eta, alpha, alpha, &t, cap);
    }
}

/* orc_ggx_ms_eval_batch on the Beckmann walk. */
void orc_beckmann_ms_eval_batch(int n, const float* wi, const float* wo, float eta, float alpha, uint32_t seed, int cap,
                                float* out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        uint32_t t = orc_tea16((uint32_t)i, seed);
        out[i] = (float)orc_beckmann_ms_dielectric_eval(V3(wi[0], wi[1], wi[2]),
                                                        V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), eta, alpha, alpha,
                                                        &t, cap);
    }
}

/* Rough diffuse walks (dist 1 GGX, 0 Beckmann) from the fixed local wi, albedo
 * rho[3] (stream tea16(i, seed)): exit direction, last visible normal, weight
 * (3 floats each) and order. */
void orc_ms_diffuse_sample_batch(int n, const float* wi, float alpha, int dist, const float* rho, uint32_t seed, int cap,
                                 float* wo, float* wm, float* weight, int32_t* order)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        uint32_t t = orc_tea16((uint32_t)i, seed);
        v3 o, m, w;
        int k = 0;
        orc_ms_diffuse_sample(dist, V3(wi[0], wi[1], wi[2]), alpha, alpha, V3(rho[0], rho[1], rho[2]), &t, cap, &o, &m,
                              &w, &k);
        order[i] = k;
        const v3 r[3] = {o, m, w};
        float* dst[3] = {wo, wm, weight};
        for (int j = 0; j < 3; ++j) {
            dst[j][3 * i] = (float)r[j].x;
            dst[j][3 * i + 1] = (float)r[j].y;
            dst[j][3 * i + 2] = (float)r[j].z;
        }
    }
}

/* One rough diffuse walk estimate of BSDF x cos (3 channels) per direction wo[i]
 * for the fixed local wi (stream tea16(i, seed)), at most `cap` events. */
void orc_ms_diffuse_eval_batch(int n, const float* wi, const float* wo, float alpha, int dist, const float* rho,
                               uint32_t seed, int cap, float* out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        uint32_t t = orc_tea16((uint32_t)i, seed);
        v3 r = orc_ms_diffuse_eval(dist, V3(wi[0], wi[1], wi[2]), V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), alpha,
                                   alpha, V3(rho[0], rho[1], rho[2]), &t, cap);
        out[3 * i] = (float)r.x;
        out[3 * i + 1] = (float)r.y;
        out[3 * i + 2] = (float)r.z;
    }
}

/* One conductor walk estimate of BSDF x cos per direction wo[i] for the fixed
 * local wi, complex ior per channel as 6 doubles (eta = 1 / ior). */
void orc_ms_conductor_eval_batch(int n, const float* wi, const float* wo, const double* ior, float alpha, int dist,
                                 uint32_t seed, int cap, float* out)
{
    cplx3 eta, ieta;
    ieta.x = (cplx){ior[0], ior[1]};
    ieta.y = (cplx){ior[2], ior[3]};
    ieta.z = (cplx){ior[4], ior[5]};
    eta.x = orc_cdiv_fc((R)1.0f, ieta.x);
    eta.y = orc_cdiv_fc((R)1.0f, ieta.y);
    eta.z = orc_cdiv_fc((R)1.0f, ieta.z);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        uint32_t t = orc_tea16((uint32_t)i, seed);
        v3 r = orc_ms_conductor_eval(dist, V3(wi[0], wi[1], wi[2]), V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), &eta,
                                     &ieta, alpha, alpha, &t, cap);
        out[3 * i] = (float)r.x;
        out[3 * i + 1] = (float)r.y;
        out[3 * i + 2] = (float)r.z;
    }
}

/* Walter-model microfacet normals about +z (microfacet_sample_normal). */
void orc_walter_sample_batch(int n, const float* z1, const float* z2, float alpha, int dist, float* m)
{
    for (int i = 0; i < n; i++) {
        v3 r = orc_sample_normal_walter(V3(0, 0, 1), z1[i], z2[i], alpha, dist);
        m[3 * i] = (float)r.x;
        m[3 * i + 1] = (float)r.y;
        m[3 * i + 2] = (float)r.z;
    }
}

/* Reflection weight f cos / pdf with F = 1 for incidence wi (local, +z normal)
 * under the Walter model (normal ~ D cos, weight |i.m| G1(i) G1(o) / (|i.n|
 * |n.m|)) or visible normals (weight G1(o)); 0 where the shaders stop the path
 * (G1 <= 0).  Both estimate the same directional albedo. */
void orc_reflect_weight_batch(int n, const float* wi, const float* z1, const float* z2, float alpha, int dist,
                              int model, float* out)
{
    v3 w_i = V3(wi[0], wi[1], wi[2]), nz = V3(0, 0, 1);
    for (int i = 0; i < n; i++) {
        R w = 0;
        if (model == 0) {
            v3 m = orc_sample_normal_walter(nz, z1[i], z2[i], alpha, dist);
            R gi = orc_masking_iso(w_i, m, nz, alpha, dist);
            v3 wo = vadd(vneg(w_i), vscale(vscale(m, (R)2.0f), vdot(w_i, m)));
            R go = orc_masking_iso(wo, m, nz, alpha, dist);
            if (gi > 0 && go > 0) w = fabs(vdot(w_i, m)) * gi * go / (fabs(vdot(w_i, nz)) * fabs(vdot(nz, m)));
        } else {
            v3 l = w_i;
            v3 m = dist == 1 ? orc_ggx_sample_vndf(l, alpha, alpha, z1[i], z2[i])
                             : orc_beckmann_sample_vndf(l, alpha, alpha, z1[i], z2[i]);
            R gi = dist == 1 ? orc_ggx_G1(w_i, m, nz, alpha, alpha) : orc_beckmann_G1(w_i, m, nz, alpha, alpha);
            v3 wo = vadd(vneg(w_i), vscale(vscale(m, (R)2.0f), vdot(w_i, m)));
            R go = dist == 1 ? orc_ggx_G1(wo, m, nz, alpha, alpha) : orc_beckmann_G1(wo, m, nz, alpha, alpha);
            if (gi > 0 && go > 0) w = go;
        }
        out[i] = (float)w;
    }
}
