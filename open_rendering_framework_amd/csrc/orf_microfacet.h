// orf_microfacet.h -- GGX microfacet model, frames and samplers (device and host).
//
// Formulas and their evaluation order follow the reference headers so results
// agree to rounding with the CPU restatement:
//   frames           ORF/helpers.h:107-119 (create_onb), :284-299 (to local/world)
//   samplers         ORF/sampler.h:12-24 (rotate_to_normal), :51-64 (cosine), :103-116 (barycentric)
//   GGX              ORF/MicrofacetGGX.h:9-17 (alpha_i), :33-45 (projected area),
//                    :48-60 (slope pdf, Smith G1), :63-74 (G1 with frame),
//                    :142-187 (P22_11), :190-207 (VNDF sampling), :219-251 (NDF, VNDF eval)
// Transcendentals are the shared FP32 functions of orf_math.h (the same bits on
// the GPU and in the CPU oracle); the identity cos(acos x) is deliberately NOT
// simplified: the reference rounds through theta, and so do these kernels.
#pragma once

#include "orf_vec.h"
#include "orf_rng.h"

namespace orf {

// Orthonormal frame around a unit normal n; returns (u, v) with v = n x u.
ORF_HD void make_frame(f3 n, f3& u, f3& v)
{
    u = cross(n, mk3(0.0f, 1.0f, 0.0f));
    if (dot(u, u) < 1e-3f) u = cross(n, mk3(1.0f, 0.0f, 0.0f));
    u = normalize(u);
    v = cross(n, u);
}

ORF_HD f3 to_local(f3 w, f3 n)
{
    f3 u, v;
    make_frame(n, u, v);
    return normalize(mk3(dot(w, u), dot(w, v), dot(w, n)));
}

ORF_HD f3 to_world(f3 w, f3 n)
{
    f3 u, v;
    make_frame(n, u, v);
    return normalize(u * w.x + v * w.y + n * w.z);
}

// Rotation taking +z to n applied to v (Frisvad 2012, branch-free ONB).
ORF_HD f3 rotate_to_normal(f3 n, f3 v)
{
    if (n.z < -0.999999f) return mk3(-v.y, -v.x, -v.z);
    float a = 1.0f / (1.0f + n.z);
    float b = -n.x * n.y * a;
    return mk3(1.0f - n.x * n.x * a, b, -n.x) * v.x + mk3(b, 1.0f - n.y * n.y * a, -n.y) * v.y + n * v.z;
}

ORF_HD f3 sample_cosine(f3 n, uint32_t& rng)
{
    float ct = sqrtf(rnd_tea(rng));
    float phi = 2.0f * kPi * rnd_tea(rng);
    float st = sqrtf(1.0f - ct * ct);
    float sp, cp;
    orf_sincosf(phi, &sp, &cp);
    return rotate_to_normal(n, mk3(st * cp, st * sp, ct));
}

ORF_HD f3 sample_barycentric(uint32_t& rng)
{
    float s = sqrtf(rnd_tea(rng));
    float xi2 = rnd_tea(rng);
    return mk3(1.0f - s, (1.0f - xi2) * s, xi2 * s);
}

// ---------------------------------------------------------------- GGX

// Roughness projected on the azimuth of wi (wi local, n = +z).
ORF_HD float ggx_alpha_proj(f3 wi, float ax, float ay)
{
    float inv_s2 = 1.0f / (1.0f - wi.z * wi.z);
    if (wi.z == 1.0f || wi.z == -1.0f) inv_s2 = 0.0f;
    float c2 = wi.x * wi.x * inv_s2;
    float s2 = wi.y * wi.y * inv_s2;
    return sqrtf(c2 * ax * ax + s2 * ay * ay);
}

// Smith masking for GGX from cos(theta) and alpha^2: 2 / (1 + sqrt(1 + a^2 tan^2)).
ORF_HD float ggx_smith_g1(float cos_theta, float a2)
{
    float c2 = cos_theta * cos_theta;
    float t2 = (1.0f - c2) / c2;
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * t2));
}

// Masking of direction v (world) on microfacet m over macro-normal n.
ORF_HD float ggx_masking(f3 v, f3 m, f3 n, float ax, float ay)
{
    float a = ggx_alpha_proj(normalize(to_local(v, n)), ax, ay);
    float a2 = a * a;
    float cv = dot(v, n);
    if (cv * dot(v, m) <= 0.0f) return 0.0f;
    return ggx_smith_g1(cv, a2);
}

// cos(theta)(1 + Lambda(wi)): the normaliser of the visible-normal density.
ORF_HD float ggx_proj_area(f3 wi, float ax, float ay)
{
    if (wi.z > 0.9999f) return 1.0f;
    if (wi.z < -0.9999f) return 0.0f;
    float a = ggx_alpha_proj(wi, ax, ay);
    float s = orf_sinf(orf_acosf(wi.z));
    return 0.5f * (wi.z + sqrtf(wi.z * wi.z + s * s * a * a));
}

ORF_HD float ggx_ndf(f3 wm, float ax, float ay)
{
    if (wm.z <= 0.0f) return 0.0f;
    float sx = -wm.x / wm.z;
    float sy = -wm.y / wm.z;
    float t = 1.0f + sx * sx / (ax * ax) + sy * sy / (ay * ay);
    float p22 = 1.0f / (kPi * ax * ay) / (t * t);
    return p22 / (wm.z * wm.z * wm.z * wm.z);
}

// Density of visible normals D_wi(m) = <wi,m>+ D(m) / proj_area(wi).
ORF_HD float ggx_vndf(f3 wi, f3 wm, float ax, float ay)
{
    if (wm.z <= 0.0f) return 0.0f;
    float area = ggx_proj_area(wi, ax, ay);
    if (area == 0.0f) return 0.0f;
    float c = 1.0f / area;
    float D = ggx_ndf(wm, ax, ay);
    return c * fmaxf(0.0f, dot(wi, wm)) * D;
}

// Visible slopes of the unit-roughness GGX for incidence theta_i
// (Heitz & d'Eon 2014 inversion with their rational fit for the y slope).
ORF_HD f2 ggx_visible_slope(float theta_i, float z1, float z2)
{
    if (theta_i < 0.0001f) {
        float r = sqrtf(z1 / (1.0f - z1));
        float sp, cp;
        orf_sincosf(kPi * 2.0f * z2, &sp, &cp);
        return mk2(r * cp, r * sp);
    }
    float sin_i, cos_i;
    orf_sincosf(theta_i, &sin_i, &cos_i);
    cos_i = fmaxf(cos_i, 1e-7f);
    float tan_i = sin_i / cos_i;
    float proj = 0.5f * (cos_i + 1.0f);
    if (proj < 0.0001f || proj != proj) return mk2(0.0f, 0.0f);
    float c = 1.0f / proj;
    float A = 2.0f * z1 / cos_i / c - 1.0f;
    float B = tan_i;
    float tmp = 1.0f / (A * A - 1.0f);
    float disc = B * B * tmp * tmp - (A * A - B * B) * tmp;
    float D = sqrtf(disc > 0.0f ? disc : 0.0f);
    float sx1 = B * tmp - D;
    float sx2 = B * tmp + D;
    f2 s;
    s.x = (A < 0.0f || sx2 > 1.0f / tan_i) ? sx1 : sx2;
    float S, zz2;
    if (z2 > 0.5f) {
        S = 1.0f;
        zz2 = 2.0f * (z2 - 0.5f);
    } else {
        S = -1.0f;
        zz2 = 2.0f * (0.5f - z2);
    }
    float z = (zz2 * (zz2 * (zz2 * 0.27385f - 0.73369f) + 0.46341f)) /
              (zz2 * (zz2 * (zz2 * 0.093073f + 0.309420f) - 1.0000f) + 0.597999f);
    s.y = S * z * sqrtf(1.0f + s.x * s.x);
    return s;
}

// Sample a visible microfacet normal for local wi (n = +z).
ORF_HD f3 ggx_sample_visible(f3 wi, float ax, float ay, float z1, float z2)
{
    f3 w11 = normalize(mk3(ax * wi.x, ay * wi.y, wi.z));
    f2 s11 = ggx_visible_slope(orf_acosf(w11.z), z1, z2);
    float phi = orf_atan2f(w11.y, w11.x);
    float sp, cp;
    orf_sincosf(phi, &sp, &cp);
    float sx = cp * s11.x - sp * s11.y;
    float sy = sp * s11.x + cp * s11.y;
    sx *= ax;
    sy *= ay;
    if (!__builtin_isfinite(sx)) {
        if (wi.z > 0.0f) return mk3(0.0f, 0.0f, 1.0f);
        return normalize(mk3(wi.x, wi.y, 0.0f));
    }
    return normalize(mk3(-sx, -sy, 1.0f));
}

// ---------------------------------------------------------------- Beckmann
//
// ORF/MicrofacetBeckmann.h:13-21 (alpha_i), :68-98 (G1 rational fit), :161-210
// (visible-slope sampling by Newton-bisection in erf space), :213-231 (VNDF).
// The reference evaluates erfinv in FP64 (Giles 2010 double-precision
// polynomial, ORF/helpers.h:301-377) and a few sub-expressions in double; on the
// GPU everything is FP32 with Giles' single-precision erfinv polynomial from the
// same paper (max relative error 2.4e-7 against a library erfinv).

constexpr float k2SqrtPi = 1.12837916709551257390f;

ORF_HD float erfinv_f32(float x)
{
    float w = -orf_logf((1.0f - x) * (1.0f + x)), p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = 3.43273939e-07f + p * w;
        p = -3.5233877e-06f + p * w;
        p = -4.39150654e-06f + p * w;
        p = 0.00021858087f + p * w;
        p = -0.00125372503f + p * w;
        p = -0.00417768164f + p * w;
        p = 0.246640727f + p * w;
        p = 1.50140941f + p * w;
    } else {
        w = sqrtf(w) - 3.0f;
        p = -0.000200214257f;
        p = 0.000100950558f + p * w;
        p = 0.00134934322f + p * w;
        p = -0.00367342844f + p * w;
        p = 0.00573950773f + p * w;
        p = -0.0076224613f + p * w;
        p = 0.00943887047f + p * w;
        p = 1.00167406f + p * w;
        p = 2.83297682f + p * w;
    }
    return p * x;
}

ORF_HD float beckmann_smith_g1(float cos_theta, float a2)
{
    float c2 = cos_theta * cos_theta;
    float t2 = (1.0f - c2) / c2;
    float a_sqr = 1.0f / (a2 * t2);
    float a = sqrtf(a_sqr);
    return a < 1.6f ? (3.535f * a + 2.181f * a_sqr) / (1.0f + 2.276f * a + 2.577f * a_sqr) : 1.0f;
}

ORF_HD float beckmann_masking(f3 v, f3 m, f3 n, float ax, float ay)
{
    float width = ggx_alpha_proj(to_local(v, n), ax, ay);
    float cv = dot(v, n);
    if (dot(v, m) * cv <= 0.0f) return 0.0f;
    return beckmann_smith_g1(cv, width * width);
}

ORF_HD f2 beckmann_visible_slope(float theta_i, float z1, float z2)
{
    if (theta_i < 0.0001f) {
        float r = sqrtf(-orf_logf(z1));
        float sp, cp;
        orf_sincosf(kPi * 2.0f * z2, &sp, &cp);
        return mk2(r * cp, r * sp);
    }
    float sin_i, cos_i;
    orf_sincosf(theta_i, &sin_i, &cos_i);
    float slope_i = cos_i / sin_i;
    float a = cos_i / sin_i;
    float proj = 0.5f * (orf_erff(a) + 1.0f) * cos_i + 0.25f * k2SqrtPi * sin_i * orf_expf(-a * a);
    if (proj < 0.0001f || proj != proj) return mk2(0.0f, 0.0f);
    float c = 1.0f / proj;
    float emin = -0.9999f;
    float emax = fmaxf(emin, orf_erff(slope_i));
    float ecur = 0.5f * (emin + emax);
    for (int it = 0; it < 256 && emax - emin > 0.00001f; ++it) {
        if (!(ecur >= emin && ecur <= emax)) ecur = 0.5f * (emin + emax);
        float s = erfinv_f32(ecur);
        float cdf = (s >= slope_i) ? 1.0f
                                   : c * (0.25f * k2SqrtPi * sin_i * orf_expf(-s * s) +
                                          cos_i * (0.5f + 0.5f * orf_erff(s)));
        float diff = cdf - z1;
        if (fabsf(diff) < 0.00001f) break;
        if (diff > 0.0f) {
            if (emax == ecur) break;
            emax = ecur;
        } else {
            if (emin == ecur) break;
            emin = ecur;
        }
        float deriv = 0.5f * c * cos_i - 0.5f * c * sin_i * s;
        ecur -= diff / deriv;
    }
    return mk2(erfinv_f32(fminf(emax, fmaxf(emin, ecur))), erfinv_f32(2.0f * z2 - 1.0f));
}

ORF_HD f3 beckmann_sample_visible(f3 wi, float ax, float ay, float z1, float z2)
{
    f3 w11 = normalize(mk3(ax * wi.x, ay * wi.y, wi.z));
    f2 s11 = beckmann_visible_slope(orf_acosf(w11.z), z1, z2);
    float phi = orf_atan2f(w11.y, w11.x);
    float sp, cp;
    orf_sincosf(phi, &sp, &cp);
    float sx = cp * s11.x - sp * s11.y;
    float sy = sp * s11.x + cp * s11.y;
    sx *= ax;
    sy *= ay;
    if (!__builtin_isfinite(sx)) {
        if (wi.z > 0.0f) return mk3(0.0f, 0.0f, 1.0f);
        return normalize(mk3(wi.x, wi.y, 0.0f));
    }
    return normalize(mk3(-sx, -sy, 1.0f));
}

// microfacet_sample_visible_normal / masking_G1 dispatch (ORF/Microfacet.h:24-32, 72-83).
ORF_HD f3 sample_visible_normal(f3 wi, f3 n, float ax, float ay, float z1, float z2, int dist)
{
    f3 l = to_local(wi, n);
    f3 m = dist == 1 ? ggx_sample_visible(l, ax, ay, z1, z2) : beckmann_sample_visible(l, ax, ay, z1, z2);
    return to_world(m, n);
}

ORF_HD float masking_g1(f3 v, f3 m, f3 n, float ax, float ay, int dist)
{
    return dist == 1 ? ggx_masking(v, m, n, ax, ay) : beckmann_masking(v, m, n, ax, ay);
}

// ---------------------------------------------------------------- Walter et al. 2007 model

// Isotropic masking with roughness a: ggx_G1(v, m, n, a_g) / beckmann_G1(v, m,
// n, width) (ORF/MicrofacetGGX.h:77-83, ORF/MicrofacetBeckmann.h:78-86), the
// form the reference's Walter-model branches call (Microfacet.h:16-22).
ORF_HD float masking_g1_iso(f3 v, f3 m, f3 n, float a, int dist)
{
    float cv = dot(v, n);
    if (dist == 1) {
        if (cv * dot(v, m) <= 0.0f) return 0.0f;
        return ggx_smith_g1(cv, a * a);
    }
    if (dot(v, m) * cv <= 0.0f) return 0.0f;
    return beckmann_smith_g1(cv, a * a);
}

// microfacet_sample_normal (ORF/Microfacet.h:50-60) -> ggx_sample_hemisphere /
// beckmann_sample_hemisphere (ORF/MicrofacetGGX.h:123-139,
// ORF/MicrofacetBeckmann.h:144-158): a normal with density D(m) cos(theta_m)
// about n (tan^2 theta = a^2 z1 / (1 - z1) for GGX, -a^2 log(1 - z1) for
// Beckmann, phi = 2 pi z2), placed with the create_onb frame, not renormalised.
ORF_HD f3 sample_normal_walter(f3 n, float z1, float z2, float a, int dist)
{
    float phi = 2.0f * kPi * z2;
    float theta;
    if (dist == 1) {
        float tan_theta = a * sqrtf(z1) / sqrtf(1.0f - z1);
        theta = orf_atanf(tan_theta);
    } else {
        float tan_theta_sqr = -a * a * orf_logf(1.0f - z1);
        theta = orf_atanf(sqrtf(tan_theta_sqr));
    }
    float ct = orf_cosf(theta), st = orf_sinf(theta);
    f3 l = mk3(st * orf_cosf(phi), st * orf_sinf(phi), ct);
    f3 u, v;
    make_frame(n, u, v);
    return u * l.x + v * l.y + n * l.z;
}

// Walter-model weight f cos / pdf of a direction scattered off the sampled
// normal m: |i.m| G1(i) G1(o) / (|i.n| |n.m|) (CU/metallic_shader.cu:148,
// CU/rough_transparent_shader.cu:184-200).
ORF_HD float walter_weight(f3 w_i, f3 m, f3 ffn, f3 normal, float g_i, float g_o)
{
    float abs_i_m = fabsf(dot(w_i, m));
    float abs_i_n = fabsf(dot(w_i, ffn));
    float abs_n_m = fabsf(dot(normal, m));
    return abs_i_m * g_i * g_o / (abs_i_n * abs_n_m);
}

// ---------------------------------------------------------------- Beckmann evaluation

// beckmann_projected_area (ORF/MicrofacetBeckmann.h:45-58); wi local.
ORF_HD float beckmann_proj_area(f3 wi, float ax, float ay)
{
    if (wi.z > 0.9999f) return 1.0f;
    if (wi.z < -0.9999f) return 0.0f;
    float alpha = ggx_alpha_proj(wi, ax, ay);  // beckmann_alpha_i is the same expression
    float theta = orf_acosf(wi.z);
    float a = 1.0f / orf_tanf(theta) / alpha;
    return 0.5f * (orf_erff(a) + 1.0f) * wi.z + 0.25f * k2SqrtPi * alpha * orf_sinf(theta) * orf_expf(-a * a);
}

// beckmann_slope_pdf / beckmann_eval_NDF (ORF/MicrofacetBeckmann.h:61-65, 249-261).
ORF_HD float beckmann_ndf(f3 wm, float ax, float ay)
{
    if (wm.z <= 0.0f) return 0.0f;
    float sx = -wm.x / wm.z, sy = -wm.y / wm.z;
    float p22 = 1.0f / (kPi * ax * ay) * orf_expf(-sx * sx / (ax * ax) - sy * sy / (ay * ay));
    return p22 / (wm.z * wm.z * wm.z * wm.z);
}

// beckmann_eval_VNDF (ORF/MicrofacetBeckmann.h:264-281).
ORF_HD float beckmann_vndf(f3 wi, f3 wm, float ax, float ay)
{
    if (wm.z <= 0.0f) return 0.0f;
    float area = beckmann_proj_area(wi, ax, ay);
    if (area == 0.0f) return 0.0f;
    float c = 1.0f / area;
    return c * fmaxf(0.0f, dot(wi, wm)) * beckmann_ndf(wm, ax, ay);
}

// microfacet_eval_visible_normal (ORF/Microfacet.h:85-97): D_wi(wm) about n.
ORF_HD float eval_visible_normal(f3 wi, f3 wm, f3 n, float ax, float ay, int dist)
{
    f3 li = to_local(wi, n), lm = to_local(wm, n);
    return dist == 1 ? ggx_vndf(li, lm, ax, ay) : beckmann_vndf(li, lm, ax, ay);
}

}  // namespace orf
