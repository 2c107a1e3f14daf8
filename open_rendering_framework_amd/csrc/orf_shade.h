// orf_shade.h -- environment, light sampling and per-material shading steps (device).
//
// Environment:  env_lookup follows ORF/Envmap.h:57-64 (lat-long u = (atan2(x,z)+pi)/2pi,
//   v = 1 - acos(y)/pi) with an exact FP32 bilinear fetch, wrap = repeat on both axes
//   (ORF/HDRLoader.cpp:231-286), texel centres at (i + 0.5) / w.
//   env_sample follows sample_environment, ORF/Envmap.h:66-98 (two draws, marginal then
//   conditional inverse CDF with linear remap, L = E sin(theta) 2 pi^2 / p_uv); the CDF
//   search is a lower bound (first index with cdf >= xi).  env_pdf is the matching
//   solid-angle density p_uv / (2 pi^2 sin(theta)).
// Lights: ORF/LightSampler.h:20-37 (point, directional), :40-71 (triangle area light).
// Lambert step: CU/diffuse_shader.cu:42-107; conductor step: CU/metallic_shader.cu:84-158.
// The environment as a light (NEE + power-heuristic MIS against cosine sampling) is
// new: the reference's env importance sampling (CU/env_cameras.cu, CU/envmap_shader.cu)
// is never compiled and its miss program ignores emit_light.
#pragma once

#include "orf_fresnel.h"
#include "orf_microfacet.h"
#include "orf_rng.h"
#include "orf_scene.h"

namespace orf {

ORF_HD f3 tex_bilinear(const float4* tex, int w, int h, float u, float v)
{
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float fx = x - fx0, fy = y - fy0;
    int ix = (int)fx0, iy = (int)fy0;
    ix %= w;
    if (ix < 0) ix += w;
    iy %= h;
    if (iy < 0) iy += h;
    int ix1 = ix + 1 == w ? 0 : ix + 1;
    int iy1 = iy + 1 == h ? 0 : iy + 1;
    float4 t00 = tex[(size_t)iy * w + ix], t10 = tex[(size_t)iy * w + ix1];
    float4 t01 = tex[(size_t)iy1 * w + ix], t11 = tex[(size_t)iy1 * w + ix1];
    float gx = 1.0f - fx, gy = 1.0f - fy;
    f3 r0 = mk3(t00.x * gx + t10.x * fx, t00.y * gx + t10.y * fx, t00.z * gx + t10.z * fx);
    f3 r1 = mk3(t01.x * gx + t11.x * fx, t01.y * gx + t11.y * fx, t01.z * gx + t11.z * fx);
    return mk3(r0.x * gy + r1.x * fy, r0.y * gy + r1.y * fy, r0.z * gy + r1.z * fy);
}

ORF_HD f3 env_lookup(const EnvGPU& env, f3 dir)
{
    float theta = orf_acosf(dir.y);
    float phi = orf_atan2f(dir.x, dir.z);
    float u = (phi + kPi) * 0.5f * kInvPi;
    float v = (float)(1.0 - (double)(theta * kInvPi));
    return tex_bilinear(env.tex, env.w, env.h, u, v);
}

// First index i in [0, n) with cdf[i] > xi (n - 1 if none): the cell of xi,
// never one of zero width (the reference's bisection, ORF/Envmap.h:23-55, can
// return a zero-width cell when xi equals a flat run's value, e.g. xi = 0 with
// zero-density leading cells, and the sampler then divides 0/0).
ORF_HD int cdf_upper_bound(const float* cdf, int n, float xi)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] > xi) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Solid-angle density of env_sample for direction dir.
ORF_HD float env_pdf(const EnvGPU& env, f3 dir)
{
    float theta = orf_acosf(dir.y);
    float phi = orf_atan2f(dir.x, dir.z);
    float u = (phi + kPi) * 0.5f * kInvPi;
    float v = 1.0f - theta * kInvPi;
    int iu = min(env.w - 1, max(0, (int)(u * (float)env.w)));
    int iv = min(env.h - 1, max(0, (int)(v * (float)env.h)));
    float p = env.marg_pdf[iv] * env.cond_pdf[(size_t)iv * env.w + iu];
    float st = orf_sinf(theta);
    return st > 0.0f ? p / (2.0f * kPi * kPi * st) : 0.0f;
}

// Sample a direction from the environment; L = radiance / pdf (solid angle).
ORF_HD void env_sample(const EnvGPU& env, uint32_t& rng, f3& dir, f3& L, float& pdf_sa)
{
    const float two_pi2 = 2.0f * kPi * kPi;
    float xi1 = rnd_tea(rng), xi2 = rnd_tea(rng);
    int vi = cdf_upper_bound(env.marg_cdf, env.h, xi1);
    float c0 = vi > 0 ? env.marg_cdf[vi - 1] : 0.0f;
    float dv = vi > 0 ? (xi1 - c0) / (env.marg_cdf[vi] - c0) : xi1 / env.marg_cdf[vi];
    float pdf_m = env.marg_pdf[vi];
    float v = ((float)vi + dv) / (float)env.h;
    const float* ccdf = env.cond_cdf + (size_t)vi * env.w;
    int ui = cdf_upper_bound(ccdf, env.w, xi2);
    float c1 = ui > 0 ? ccdf[ui - 1] : 0.0f;
    float du = ui > 0 ? (xi2 - c1) / (ccdf[ui] - c1) : xi2 / ccdf[ui];
    float pdf_c = env.cond_pdf[(size_t)vi * env.w + ui];
    float u = ((float)ui + du) / (float)env.w;
    float prob = pdf_m * pdf_c;
    float theta = (1.0f - v) * kPi;
    float phi = (2.0f * u - 1.0f) * kPi;
    float st, ct, sp, cp;
    orf_sincosf(theta, &st, &ct);
    orf_sincosf(phi, &sp, &cp);
    dir = mk3(st * sp, ct, st * cp);
    f3 E = tex_bilinear(env.tex, env.w, env.h, u, v);
    L = E * (st * two_pi2 / prob);
    // a draw that rounds onto a pole (st = 0) carries L = 0; its density is
    // reported as 0 (not +inf) so the MIS weight stays finite
    pdf_sa = st > 0.0f ? prob / (two_pi2 * st) : 0.0f;
}

// evaluate_direct_illumination (ORF/LightSampler.h:143-166) for one picked light.
// Returns false for light types without a sampler here.  kAnalytic: disk and
// sphere lights are sampled (kernel variants for scenes without them leave
// that code out).
template <bool kAnalytic = true>
ORF_HD bool sample_light(const SceneGPU& sc, const LightGPU& lt, f3 pos, uint32_t& rng, f3& dir, f3& L,
                         float& dist, float& pdf_env)
{
    pdf_env = -1.0f;
    if (lt.type == POINT_LIGHT) {
        dir = lt.params - pos;
        float r2 = dot(dir, dir);
        dist = sqrtf(r2);
        dir = dir / dist;
        L = lt.radiance / r2;
        return true;
    }
    if (lt.type == DIRECTIONAL_LIGHT) {
        dir = normalize(-lt.params);
        dist = kRayMax;
        L = lt.radiance;
        return true;
    }
    if (lt.type == TRIANGLES_AREA_LIGHT) {
        unsigned ntri = (unsigned)lt.params.x;
        unsigned pick = (unsigned)(rnd_tea(rng) * (float)ntri);
        if (pick >= ntri) pick = ntri - 1;  // rnd_tea can round to 1.0f
        const TriLightGPU& tl = sc.tri_lights[pick + (unsigned)lt.params.y];
        f3 uvw = sample_barycentric(rng);
        f3 lp = tl.v0 * uvw.x + tl.v1 * uvw.y + tl.v2 * uvw.z;
        f3 n = tl.has_normals ? normalize(tl.n0 * uvw.x + tl.n1 * uvw.y + tl.n2 * uvw.z)
                              : normalize(cross(tl.v1 - tl.v0, tl.v2 - tl.v0));
        dir = lp - pos;
        float d2 = dot(dir, dir);
        dist = sqrtf(d2);
        dir = normalize(dir);
        float cp = fmaxf(dot(n, -dir), 0.0f);
        L = lt.radiance * (tl.area * cp / d2 * (float)ntri);
        return true;
    }
    if (kAnalytic && lt.type == DISK_LIGHT) {
        // evaluate_disk_area_light (ORF/LightSampler.h:73-101): a uniform point of
        // the disk (r = sqrt(u1) R, angle 2 pi u2 in the create_onb frame of its
        // normal), L = Le cos' / d^2 / (1 / area).  params: centre; pad: (radius,
        // emitter record), whose second float4 holds the normal.
        const float R = lt.pad.x;
        const float4 nb = sc.analytic[2 * (int)lt.pad.y + 1];
        const f3 n = mk3(nb.x, nb.y, nb.z);
        const float pdf = 1.0f / (kPi * R * R);
        f3 U, V;
        make_frame(n, U, V);
        float u1 = rnd_tea(rng);
        float u2 = rnd_tea(rng);
        float r = sqrtf(u1) * R;
        float t = 2.0f * kPi * u2;
        f3 dp = mk3(r * orf_cosf(t), r * orf_sinf(t), 0.0f);
        dp = dp.x * U + dp.y * V + dp.z * n;
        dir = (lt.params + dp) - pos;
        float d2 = dot(dir, dir);
        dist = sqrtf(d2);
        dir = normalize(dir);
        float cp = fmaxf(dot(n, -dir), 0.0f);
        L = lt.radiance * (cp / d2) / pdf;
        return true;
    }
    if (kAnalytic && lt.type == SPHERICAL_LIGHT) {
        // evaluate_spherical_area_light (ORF/LightSampler.h:103-141): a uniform
        // point of the hemisphere facing pos (z = u1 along the create_onb frame
        // of normalize(pos - centre)) weighted by the whole sphere's area density
        // 1 / (4 pi R^2), as the reference does (twice the hemisphere's).
        const float R = lt.pad.x;
        const f3 c = lt.params;
        const f3 vn = normalize(pos - c);
        f3 U, V;
        make_frame(vn, U, V);
        const float pdf = 1.0f / (4.0f * kPi * R * R);
        float u1 = rnd_tea(rng);
        float u2 = rnd_tea(rng);
        float z = u1;
        float s = sqrtf(1.0f - z * z);
        f3 sp = mk3(orf_cosf(2.0f * kPi * u2) * s, orf_sinf(2.0f * kPi * u2) * s, z) * R;
        sp = sp.x * U + sp.y * V + sp.z * vn + c;
        f3 n = normalize(sp - c);
        dir = sp - pos;
        float d2 = dot(dir, dir);
        dist = sqrtf(d2);
        dir = normalize(dir);
        float cp = fmaxf(dot(n, -dir), 0.0f);
        L = lt.radiance * (cp / d2) / pdf;
        return true;
    }
    if (lt.type == ENV_LIGHT) {
        env_sample(sc.env, rng, dir, L, pdf_env);
        dist = kRayMax;
        return true;
    }
    return false;
}

// One visible-normal sampling step of a rough conductor with macro-normal n
// (CU/metallic_shader.cu:84-158, VISIBLE_NORMALS_MODEL, GGX).  Returns the
// reflected direction, its solid-angle pdf and f cos / pdf (= F G1(w_o));
// the weight is 0 where the reference stops the path (G1 <= 0).
ORF_HD void ggx_conductor_sample(f3 w_i, f3 n, float ax, float ay, float z1, float z2, const ConductorEta& eta,
                                 f3& w_o, float& pdf, f3& weight)
{
    f3 m = to_world(ggx_sample_visible(to_local(w_i, n), ax, ay, z1, z2), n);
    float g_i = ggx_masking(w_i, m, n, ax, ay);
    f3 F = fresnel_conductor(w_i, m, eta);
    w_o = -w_i + (m * 2.0f) * dot(w_i, m);
    float g_o = ggx_masking(w_o, m, n, ax, ay);
    float D = ggx_vndf(to_local(w_i, n), to_local(m, n), ax, ay);
    pdf = D / (4.0f * fabsf(dot(w_o, m)));
    weight = (g_i <= 0.0f || g_o <= 0.0f) ? mk3(0.0f) : F * g_o;
}

// Conductor step for either distribution (CU/metallic_shader.cu:84-158): visible
// normals with weight F G1(w_o), or the Walter model (normal ~ D cos) with weight
// F |i.m| G1(i) G1(o) / (|i.n| |n.m|); 0 where the reference stops.
ORF_HD void conductor_step(f3 w_i, f3 n, const MaterialGPU& m, float z1, float z2, f3& w_o, f3& weight)
{
    if (m.microfacet_model == WALTER_MODEL) {
        const int dist = m.normal_distribution;
        const float a = m.roughness.x;
        f3 mn = sample_normal_walter(n, z1, z2, a, dist);
        float g_i = masking_g1_iso(w_i, mn, n, a, dist);
        w_o = -w_i + (mn * 2.0f) * dot(w_i, mn);
        if (g_i <= 0.0f) {
            weight = mk3(0.0f);
            return;
        }
        f3 F = fresnel_conductor(w_i, mn, m.eta);
        float g_o = masking_g1_iso(w_o, mn, n, a, dist);
        weight = g_o <= 0.0f ? mk3(0.0f) : F * walter_weight(w_i, mn, n, n, g_i, g_o);
        return;
    }
    if (m.normal_distribution == GGX_DISTRIBUTION) {
        float pdf;
        ggx_conductor_sample(w_i, n, m.roughness.x, m.roughness.y, z1, z2, m.eta, w_o, pdf, weight);
        return;
    }
    f3 mn = sample_visible_normal(w_i, n, m.roughness.x, m.roughness.y, z1, z2, BECKMANN_DISTRIBUTION);
    float g_i = beckmann_masking(w_i, mn, n, m.roughness.x, m.roughness.y);
    f3 F = fresnel_conductor(w_i, mn, m.eta);
    w_o = -w_i + (mn * 2.0f) * dot(w_i, mn);
    float g_o = beckmann_masking(w_o, mn, n, m.roughness.x, m.roughness.y);
    weight = (g_i <= 0.0f || g_o <= 0.0f) ? mk3(0.0f) : F * g_o;
}

// fresnel_R(cos1, cos2, ior1/ior2) (ORF/Fresnel.h:26-32).
ORF_HD float fresnel_R3(float c1, float c2, float eta)
{
    float rs = (eta * c1 - c2) / (eta * c1 + c2);
    float rp = (c1 - eta * c2) / (c1 + eta * c2);
    return (rs * rs + rp * rp) * 0.5f;
}

// fresnel_R(w_i, w_m, ior1/ior2) (ORF/Fresnel.h:42-53).
ORF_HD float fresnel_R_dir(f3 w_i, f3 w_m, float eta)
{
    float c1 = dot(w_i, w_m);
    float c22 = 1.0f - eta * eta * (1.0f - c1 * c1);
    if (c22 <= 0.0f) return 1.0f;
    return fresnel_R3(c1, sqrtf(c22), eta);
}

// Fresnel reflectance through an interface of relative index eta = n1/n2 at
// incidence cosine c; 1 under total internal reflection (the branch written out
// in CU/lambertian_interface_shader.cu:68-75).
ORF_HD float fresnel_R_cos(float c, float eta)
{
    float st2 = eta * eta * (1.0f - c * c);
    if (!(st2 < 1.0f)) return 1.0f;
    return fresnel_R3(c, sqrtf(1.0f - st2), eta);
}


// ---------------------------------------------------------------- multiple scattering
// Microsurface random walk of Heitz et al. 2016 on a Gaussian-height GGX surface
// (ORF/MicrofacetGGX.h:19-29, 93-107, 254-305, 412-438; heights ORF/helpers.h:380-391).

// ggx_smith_lambda (ORF/MicrofacetGGX.h:20-29); wi local, may point downwards.
ORF_HD float ggx_smith_lambda(f3 wi, float ax, float ay)
{
    if (wi.z > 0.99999f) return 0.0f;
    if (wi.z < -0.99999f) return -1.0f;
    float theta = orf_acosf(wi.z);
    float a = 1.0f / orf_tanf(theta) / ggx_alpha_proj(wi, ax, ay);
    return 0.5f * (-1.0f + sqrtf(1.0f + 1.0f / (a * a)) * (a > 0.0f ? 1.0f : -1.0f));
}

// GaussianHeightCDF / GaussianHeightInvCDF (ORF/helpers.h:380-391).
ORF_HD float height_cdf(float h) { return 0.5f + 0.5f * orf_erff(0.707106781186547524401f * h); }
ORF_HD float height_inv_cdf(float z) { return 1.41421356237309504880f * erfinv_f32(2.0f * z - 1.0f); }

// ggx_G(wi, h): masking of wi at height h, C1(h)^Lambda (ORF/MicrofacetGGX.h:93-107).
ORF_HD float ggx_masking_height(f3 wi, float h, float ax, float ay)
{
    if (wi.z > 0.9999f) return 1.0f;
    if (wi.z <= 0.0f) return 0.0f;
    return orf_powf(height_cdf(h), ggx_smith_lambda(wi, ax, ay));
}

// ggx_sample_height (ORF/MicrofacetGGX.h:254-275): next height along wr from hr,
// or kRayMax when the ray leaves the microsurface.
ORF_HD float ggx_sample_height(f3 wr, float hr, float z, float ax, float ay)
{
    if (wr.z > 0.9999f) return kRayMax;
    if (wr.z < -0.9999f) return height_inv_cdf(z * height_cdf(hr));
    if (fabsf(wr.z) < 0.00001f) return hr;
    float g1 = ggx_masking_height(wr, hr, ax, ay);
    if (z > 1.0f - g1) return kRayMax;
    return height_inv_cdf(height_cdf(hr) / orf_powf(1.0f - z, 1.0f / ggx_smith_lambda(wr, ax, ay)));
}

// ggx_sample_dielectric_phase (ORF/MicrofacetGGX.h:278-305): visible normal from
// the side `outside` names, Fresnel choice between reflection and refraction
// (which flips `outside`).  Draws z1, z2, then the selector.
ORF_HD f3 ggx_sample_dielectric_phase(f3 w_i, float m_eta, uint32_t& rng, bool& outside, float ax, float ay)
{
    float z1 = rnd_tea(rng);
    float z2 = rnd_tea(rng);
    bool wi_out = outside;
    float eta = wi_out ? m_eta : 1.0f / m_eta;
    f3 w_m = wi_out ? ggx_sample_visible(w_i, ax, ay, z1, z2) : -ggx_sample_visible(-w_i, ax, ay, z1, z2);
    float c = fabsf(dot(w_m, w_i));
    float st2 = eta * eta * (1.0f - c * c);
    float ct = sqrtf(1.0f - st2);
    float F = st2 < 1.0f ? fresnel_R3(c, ct, eta) : 1.0f;
    if (rnd_tea(rng) <= F) return normalize(-w_i + (w_m * 2.0f) * dot(w_i, w_m));
    outside = !wi_out;
    return normalize((-eta) * w_i + w_m * (eta * c - ct));
}

// ggx_eval_dielectric_phase (ORF/MicrofacetGGX.h:308-339): the dielectric phase
// function p(wi, wo) of the microsurface, reflection (same side) or refraction.
ORF_HD float ggx_eval_dielectric_phase(f3 w_i, f3 w_o, bool wi_out, bool wo_out, float m_eta, float ax, float ay)
{
    float eta = wi_out ? m_eta : 1.0f / m_eta;
    if (wi_out == wo_out) {
        f3 w_h = normalize(w_i + w_o);
        float D = wi_out ? ggx_vndf(w_i, w_h, ax, ay) : ggx_vndf(-w_i, -w_h, ax, ay);
        float F = wi_out ? fresnel_R_dir(w_i, w_h, eta) : fresnel_R_dir(-w_i, -w_h, eta);
        float a = wi_out ? 0.25f / dot(w_i, w_h) : 0.25f / dot(-w_i, -w_h);
        return F * D * a;
    }
    f3 w_h = -normalize(w_i + w_o / eta);
    w_h = w_h * (wi_out ? w_h.z / fabsf(w_h.z) : -w_h.z / fabsf(w_h.z));
    if (dot(w_h, w_i) < 0.0f) return 0.0f;
    float D = wi_out ? ggx_vndf(w_i, w_h, ax, ay) : ggx_vndf(-w_i, -w_h, ax, ay);
    float F = wi_out ? fresnel_R_dir(w_i, w_h, eta) : fresnel_R_dir(-w_i, -w_h, eta);
    float b = 1.0f / (eta * eta);
    float c = wi_out ? fmaxf(0.0f, -dot(w_o, w_h)) : fmaxf(0.0f, -dot(-w_o, -w_h));
    float e = wi_out ? dot(w_i, w_h) + dot(w_o, w_h) / eta : dot(-w_i, -w_h) + dot(-w_o, -w_h) / eta;
    float d = e * e;
    float a = b * c / d;
    return a * D * (1.0f - F);
}

// Scattering events allowed per walk before the path is ended (the reference's
// loop has no bound).
constexpr int kMsWalkCap = 10;

// ggx_multiscattering_dielectric_BSDF_sample (ORF/MicrofacetGGX.h:412-438): exit
// direction of the walk started by local_wi (weight 1).  Returns false when the
// walk exceeds kMsWalkCap scattering events.
ORF_HD bool ggx_ms_dielectric_sample(f3 local_wi, float eta, float ax, float ay, uint32_t& rng, f3& w_out)
{
    float hr = 1.0f + height_inv_cdf(0.9999f);
    f3 w_o = -local_wi;
    bool outside = (-w_o.z) > 0.0f;
    int order = 0;
    for (;;) {
        float z = rnd_tea(rng);
        hr = outside ? ggx_sample_height(w_o, hr, z, ax, ay) : -ggx_sample_height(-w_o, -hr, z, ax, ay);
        if (hr == kRayMax || hr == -kRayMax) break;
        if (++order > kMsWalkCap) return false;
        w_o = ggx_sample_dielectric_phase(-w_o, eta, rng, outside, ax, ay);
        if (hr != hr || w_o.z != w_o.z) {
            w_o = mk3(0.0f, 0.0f, 1.0f);
            break;
        }
    }
    w_out = w_o;
    return true;
}

// ggx_multiscattering_conductor_BSDF_sample (ORF/MicrofacetGGX.h:614-637), local
// frame: the random walk on the GGX microsurface of a conductor.  Each event
// reflects about a visible normal of the arriving direction and multiplies the
// weight by the Fresnel reflectance with eta, or 1 / eta when that direction
// points below the mean plane.  False when the walk exceeds kMsWalkCap events
// (the path ends there; the reference's loop has no bound).
ORF_HD bool ggx_ms_conductor_sample(f3 local_wi, const ConductorEta& eta, float ax, float ay, uint32_t& rng,
                                    f3& w_out, f3& weight)
{
    float hr = 1.0f + height_inv_cdf(0.9999f);
    f3 w_o = -local_wi;
    weight = mk3(1.0f);
    int order = 0;
    for (;;) {
        const float u = rnd_tea(rng);
        hr = ggx_sample_height(w_o, hr, u, ax, ay);
        if (hr == kRayMax) break;
        if (++order > kMsWalkCap) return false;
        const f3 w_t = -w_o;
        const float z1 = rnd_tea(rng);
        const float z2 = rnd_tea(rng);
        const f3 wm = ggx_sample_visible(w_t, ax, ay, z1, z2);
        w_o = -w_t + (wm * 2.0f) * dot(w_t, wm);
        weight = weight * (w_t.z > 0.0f ? fresnel_conductor(w_t, wm, eta) : fresnel_conductor_inv(w_t, wm, eta));
        if (hr != hr || w_o.z != w_o.z) {
            w_o = mk3(0.0f, 0.0f, 1.0f);
            break;
        }
    }
    w_out = w_o;
    return true;
}

// beckmann_smith_lambda (ORF/MicrofacetBeckmann.h:24-38).
ORF_HD float beckmann_smith_lambda(f3 wi, float ax, float ay)
{
    if (wi.z > 0.9999f) return 0.0f;
    if (wi.z < -0.9999f) return -1.0f;
    const float t = orf_tanf(orf_acosf(wi.z));
    if (__builtin_isinf(t)) return 0.0f;
    const float a = 1.0f / (t * ggx_alpha_proj(wi, ax, ay));  // beckmann_alpha_i is the same expression
    return 0.5f * (orf_erff(a) - 1.0f) + 0.25f * 1.12837916709551257390f / a * orf_expf(-a * a);
}

// beckmann_sample_height (ORF/MicrofacetBeckmann.h:284-306): the GGX form with
// 0.999 direction limits and the height-averaged masking 1 / (1 + Lambda)
// (beckmann_G(wi, a_x, a_y), :130-141) as the escape probability.
ORF_HD float beckmann_sample_height(f3 wr, float hr, float z, float ax, float ay)
{
    if (wr.z > 0.999f) return kRayMax;
    if (wr.z < -0.999f) return height_inv_cdf(z * height_cdf(hr));
    if (fabsf(wr.z) < 0.00001f) return hr;
    const float g1 = (wr.z > 0.9999f) ? 1.0f : (wr.z <= 0.0f ? 0.0f : 1.0f / (1.0f + beckmann_smith_lambda(wr, ax, ay)));
    if (z > 1.0f - g1) return kRayMax;
    return height_inv_cdf(height_cdf(hr) / orf_powf(1.0f - z, 1.0f / beckmann_smith_lambda(wr, ax, ay)));
}

// beckmann_sample_dielectric_phase (ORF/MicrofacetBeckmann.h:310-340): visible
// normal from the side `outside` names, reflection when the selector is <= F,
// results normalised.
ORF_HD f3 beckmann_sample_dielectric_phase(f3 w_i, float m_eta, uint32_t& rng, bool& outside, float ax, float ay)
{
    const float z1 = rnd_tea(rng);
    const float z2 = rnd_tea(rng);
    const bool wi_out = outside;
    const float eta = wi_out ? m_eta : 1.0f / m_eta;
    const f3 wm = wi_out ? beckmann_sample_visible(w_i, ax, ay, z1, z2) : -beckmann_sample_visible(-w_i, ax, ay, z1, z2);
    const float c = fabsf(dot(wm, w_i));
    const float st2 = eta * eta * (1.0f - c * c);
    const float ct = sqrtf(1.0f - st2);
    const float F = st2 < 1.0f ? fresnel_R3(c, ct, eta) : 1.0f;
    if (rnd_tea(rng) <= F) return normalize(-w_i + (wm * 2.0f) * dot(w_i, wm));
    outside = !wi_out;
    return normalize((-eta) * w_i + wm * (eta * c - ct));
}

// beckmann_multiscattering_dielectric_BSDF_sample (ORF/MicrofacetBeckmann.h:445-475).
ORF_HD bool beckmann_ms_dielectric_sample(f3 local_wi, float eta, float ax, float ay, uint32_t& rng, f3& w_out)
{
    float hr = 1.0f + height_inv_cdf(0.9999f);
    f3 w_o = -local_wi;
    bool outside = (-w_o.z) > 0.0f;
    int order = 0;
    for (;;) {
        const float z = rnd_tea(rng);
        hr = outside ? beckmann_sample_height(w_o, hr, z, ax, ay) : -beckmann_sample_height(-w_o, -hr, z, ax, ay);
        if (hr == kRayMax || hr == -kRayMax) break;
        if (++order > kMsWalkCap) return false;
        w_o = beckmann_sample_dielectric_phase(-w_o, eta, rng, outside, ax, ay);
        if (hr != hr || w_o.z != w_o.z) {
            w_o = mk3(0.0f, 0.0f, 1.0f);
            break;
        }
    }
    w_out = w_o;
    return true;
}

// beckmann_multiscattering_conductor_BSDF_sample (ORF/MicrofacetBeckmann.h:711-736).
ORF_HD bool beckmann_ms_conductor_sample(f3 local_wi, const ConductorEta& eta, float ax, float ay, uint32_t& rng,
                                         f3& w_out, f3& weight)
{
    float hr = 1.0f + height_inv_cdf(0.9999f);
    f3 w_o = -local_wi;
    weight = mk3(1.0f);
    int order = 0;
    for (;;) {
        const float u = rnd_tea(rng);
        hr = beckmann_sample_height(w_o, hr, u, ax, ay);
        if (hr == kRayMax) break;
        if (++order > kMsWalkCap) return false;
        const f3 w_t = -w_o;
        const float z1 = rnd_tea(rng);
        const float z2 = rnd_tea(rng);
        const f3 wm = beckmann_sample_visible(w_t, ax, ay, z1, z2);
        w_o = -w_t + (wm * 2.0f) * dot(w_t, wm);
        weight = weight * (w_t.z > 0.0f ? fresnel_conductor(w_t, wm, eta) : fresnel_conductor_inv(w_t, wm, eta));
        if (hr != hr || w_o.z != w_o.z) {
            w_o = mk3(0.0f, 0.0f, 1.0f);
            break;
        }
    }
    w_out = w_o;
    return true;
}

// ggx_multiscattering_dielectric_BSDF_eval (ORF/MicrofacetGGX.h:444-478): one
// random-walk estimate of the multiple-scattering BSDF times cos for (wi, wo),
// local frame: the sum over the walk's events of phase(wo) x masking of wo at
// the event's height.  Truncated after kMsWalkCap events, the orders the
// sampler keeps (the reference's loop has no bound).
ORF_HD float ggx_ms_dielectric_eval(f3 w_i, f3 w_o, float eta, float ax, float ay, uint32_t& rng)
{
    f3 w_r = -w_i;
    float hr = 1.0f + height_inv_cdf(0.9999f);
    int order = 0;
    float sum = 0.0f;
    bool wr_out = (-w_r.z) > 0.0f;
    const bool wo_out = w_o.z > 0.0f;
    for (;;) {
        float z = rnd_tea(rng);
        hr = wr_out ? ggx_sample_height(w_r, hr, z, ax, ay) : -ggx_sample_height(-w_r, -hr, z, ax, ay);
        if (hr == kRayMax || hr == -kRayMax || __builtin_isinf(hr)) break;
        ++order;
        float phase = ggx_eval_dielectric_phase(-w_r, w_o, wr_out, wo_out, eta, ax, ay);
        float shadowing = wo_out ? ggx_masking_height(w_o, hr, ax, ay) : ggx_masking_height(-w_o, -hr, ax, ay);
        sum += phase * shadowing;
        w_r = ggx_sample_dielectric_phase(-w_r, eta, rng, wr_out, ax, ay);
        if (hr != hr || w_r.z != w_r.z) {
            sum = 0.0f;
            break;
        }
        if (order >= kMsWalkCap) break;
    }
    return sum;
}

// ------------------------------------------------- walks of either distribution
// The primitives above picked by the material's normal distribution, as the
// wrappers of ORF/Microfacet.h:101-184 pick them.
ORF_HD float ms_sample_height(int dist, f3 wr, float hr, float z, float ax, float ay)
{
    return dist == GGX_DISTRIBUTION ? ggx_sample_height(wr, hr, z, ax, ay) : beckmann_sample_height(wr, hr, z, ax, ay);
}

// ggx_G(wi, h) / beckmann_G(wi, h) (ORF/MicrofacetGGX.h:93-107,
// ORF/MicrofacetBeckmann.h:108-127): C1(h)^Lambda(wi), 1 near the zenith, 0 below.
ORF_HD float ms_masking_height(int dist, f3 wi, float h, float ax, float ay)
{
    if (wi.z > 0.9999f) return 1.0f;
    if (wi.z <= 0.0f) return 0.0f;
    const float lambda = dist == GGX_DISTRIBUTION ? ggx_smith_lambda(wi, ax, ay) : beckmann_smith_lambda(wi, ax, ay);
    return orf_powf(height_cdf(h), lambda);
}

ORF_HD f3 ms_sample_visible(int dist, f3 wi, float ax, float ay, float z1, float z2)
{
    return dist == GGX_DISTRIBUTION ? ggx_sample_visible(wi, ax, ay, z1, z2) : beckmann_sample_visible(wi, ax, ay, z1, z2);
}

ORF_HD float ms_vndf(int dist, f3 wi, f3 wm, float ax, float ay)
{
    return dist == GGX_DISTRIBUTION ? ggx_vndf(wi, wm, ax, ay) : beckmann_vndf(wi, wm, ax, ay);
}

// ggx_sample_diffuse_phase / beckmann_sample_diffuse_phase (ORF/MicrofacetGGX.h:
// 342-375, ORF/MicrofacetBeckmann.h:375-408): a visible normal w_m of w_i (draws
// z1, z2), then a cosine-distributed direction about it (z3, z4) through the
// concentric disk map of Shirley and Chiu 1997 in the make_frame basis of w_m,
// not renormalised.
ORF_HD f3 ms_sample_diffuse_phase(int dist, f3 w_i, float ax, float ay, uint32_t& rng, f3& w_m)
{
    const float z1 = rnd_tea(rng);
    const float z2 = rnd_tea(rng);
    const float z3 = rnd_tea(rng);
    const float z4 = rnd_tea(rng);
    w_m = ms_sample_visible(dist, w_i, ax, ay, z1, z2);
    f3 u, v;
    make_frame(w_m, u, v);
    const float r1 = 2.0f * z3 - 1.0f, r2 = 2.0f * z4 - 1.0f;
    float r, phi;
    if (r1 == 0.0f && r2 == 0.0f) {
        r = phi = 0.0f;
    } else if (r1 * r1 > r2 * r2) {
        r = r1;
        phi = (kPi / 4.0f) * (r2 / r1);
    } else {
        r = r2;
        phi = (kPi / 2.0f) - (r1 / r2) * (kPi / 4.0f);
    }
    float sp, cp;
    orf_sincosf(phi, &sp, &cp);
    const float x = r * cp, y = r * sp;
    const float z = sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
    return u * x + v * y + w_m * z;
}

// Scattering events a rough diffuse walk may take: the reference's own bound
// on its sampler (scatteringOrder < 100), also applied to the evaluation.
constexpr int kMsDiffuseCap = 100;

// ggx_multiscattering_diffuse_BSDF_sample / beckmann_... (ORF/MicrofacetGGX.h:
// 519-544, ORF/MicrofacetBeckmann.h:613-640), local frame: the walk on a
// Lambertian microsurface; every event multiplies the weight by rho.  w_m is
// the last event's visible normal (the caller's Fresnel term uses it).  After
// kMsDiffuseCap events the last sampled direction is returned as it is.
ORF_HD void ms_diffuse_sample(int dist, f3 local_wi, float ax, float ay, f3 rho, uint32_t& rng, f3& w_out, f3& w_m,
                              f3& weight)
{
    weight = mk3(1.0f);
    w_m = mk3(0.0f, 0.0f, 1.0f);
    float hr = 1.0f + height_inv_cdf(0.9999f);
    f3 w_o = -local_wi;
    for (int order = 0; order < kMsDiffuseCap;) {
        const float u = rnd_tea(rng);
        hr = ms_sample_height(dist, w_o, hr, u, ax, ay);
        if (hr == kRayMax) break;
        ++order;
        weight = weight * rho;
        w_o = ms_sample_diffuse_phase(dist, -w_o, ax, ay, rng, w_m);
        if (hr != hr || w_o.z != w_o.z) {
            w_o = mk3(0.0f, 0.0f, 1.0f);
            break;
        }
    }
    w_out = w_o;
}

// ggx_multiscattering_diffuse_BSDF_eval / beckmann_... (ORF/MicrofacetGGX.h:
// 547-577, ORF/MicrofacetBeckmann.h:643-674): one walk estimate of BSDF x cos
// for (w_i, w_o), the sum over the events of rho (once, as written there) x the
// diffuse phase about a freshly drawn visible normal (max(0, w_o.w_m) / pi) x
// the masking of w_o at the event's height; non-finite terms are skipped.
// Truncated after kMsDiffuseCap events.
ORF_HD f3 ms_diffuse_eval(int dist, f3 w_i, f3 w_o, float ax, float ay, f3 rho, uint32_t& rng)
{
    f3 w_r = -w_i;
    float hr = 1.0f + height_inv_cdf(0.9999f);
    f3 sum = mk3(0.0f);
    for (int order = 0; order < kMsDiffuseCap; ++order) {
        const float u = rnd_tea(rng);
        hr = ms_sample_height(dist, w_r, hr, u, ax, ay);
        if (hr == kRayMax) break;
        const float z1 = rnd_tea(rng);
        const float z2 = rnd_tea(rng);
        const f3 wm = ms_sample_visible(dist, -w_r, ax, ay, z1, z2);
        const float phase = kInvPi * fmaxf(0.0f, dot(w_o, wm));
        const f3 I = rho * (phase * ms_masking_height(dist, w_o, hr, ax, ay));
        if (__builtin_isfinite(I.x) && __builtin_isfinite(I.y) && __builtin_isfinite(I.z)) sum = sum + I;
        f3 wm_s;
        w_r = ms_sample_diffuse_phase(dist, -w_r, ax, ay, rng, wm_s);
        if (hr != hr || w_r.z != w_r.z) return mk3(0.0f);
    }
    return sum;
}

// ggx_multiscattering_conductor_BSDF_eval / beckmann_... (ORF/MicrofacetGGX.h:
// 580-611, ORF/MicrofacetBeckmann.h:677-708): the sum over the walk's events of
// F(-w_r, h) x 0.25 D_{-w_r}(h) (h the half vector of -w_r and w_o, 0 when it
// points below the mean plane) x the masking of w_o at the event's height, with
// eta or 1 / eta as the arriving direction lies above or below; non-finite
// terms are skipped.  Truncated after kMsWalkCap events, the orders the
// conductor sampler keeps.
ORF_HD f3 ms_conductor_eval(int dist, f3 w_i, f3 w_o, const ConductorEta& eta, float ax, float ay, uint32_t& rng)
{
    f3 w_r = -w_i;
    float hr = 1.0f + height_inv_cdf(0.9999f);
    f3 sum = mk3(0.0f);
    for (int order = 0; order < kMsWalkCap; ++order) {
        const float u = rnd_tea(rng);
        hr = ms_sample_height(dist, w_r, hr, u, ax, ay);
        if (hr == kRayMax) break;
        const f3 w_t = -w_r;
        const f3 h = normalize(w_t + w_o);
        const float phase = h.z < 0.0f ? 0.0f : 0.25f * ms_vndf(dist, w_t, h, ax, ay);
        const f3 F = w_t.z > 0.0f ? fresnel_conductor(w_t, h, eta) : fresnel_conductor_inv(w_t, h, eta);
        const f3 I = F * (phase * ms_masking_height(dist, w_o, hr, ax, ay));
        if (__builtin_isfinite(I.x) && __builtin_isfinite(I.y) && __builtin_isfinite(I.z)) sum = sum + I;
        const float z1 = rnd_tea(rng);
        const float z2 = rnd_tea(rng);
        const f3 wm = ms_sample_visible(dist, w_t, ax, ay, z1, z2);
        w_r = -w_t + (wm * 2.0f) * dot(w_t, wm);
        if (hr != hr || w_r.z != w_r.z) return mk3(0.0f);
    }
    return sum;
}

// beckmann_eval_dielectric_phase (ORF/MicrofacetBeckmann.h:342-372): the GGX
// form with the Beckmann visible-normal density.
ORF_HD float beckmann_eval_dielectric_phase(f3 w_i, f3 w_o, bool wi_out, bool wo_out, float m_eta, float ax, float ay)
{
    float eta = wi_out ? m_eta : 1.0f / m_eta;
    if (wi_out == wo_out) {
        f3 w_h = normalize(w_i + w_o);
        float D = wi_out ? beckmann_vndf(w_i, w_h, ax, ay) : beckmann_vndf(-w_i, -w_h, ax, ay);
        float F = wi_out ? fresnel_R_dir(w_i, w_h, eta) : fresnel_R_dir(-w_i, -w_h, eta);
        float a = wi_out ? 0.25f / dot(w_i, w_h) : 0.25f / dot(-w_i, -w_h);
        return F * D * a;
    }
    f3 w_h = -normalize(w_i + w_o / eta);
    w_h = w_h * (wi_out ? w_h.z / fabsf(w_h.z) : -w_h.z / fabsf(w_h.z));
    if (dot(w_h, w_i) < 0.0f) return 0.0f;
    float D = wi_out ? beckmann_vndf(w_i, w_h, ax, ay) : beckmann_vndf(-w_i, -w_h, ax, ay);
    float F = wi_out ? fresnel_R_dir(w_i, w_h, eta) : fresnel_R_dir(-w_i, -w_h, eta);
    float b = 1.0f / (eta * eta);
    float c = wi_out ? fmaxf(0.0f, -dot(w_o, w_h)) : fmaxf(0.0f, -dot(-w_o, -w_h));
    float e = wi_out ? dot(w_i, w_h) + dot(w_o, w_h) / eta : dot(-w_i, -w_h) + dot(-w_o, -w_h) / eta;
    return b * c / (e * e) * D * (1.0f - F);
}

// beckmann_multiscattering_dielectric_BSDF_eval (ORF/MicrofacetBeckmann.h:479-527)
// with every order summed (the render path passes order 0): the GGX estimator
// of ggx_ms_dielectric_eval on the Beckmann walk, non-finite terms skipped,
// truncated after kMsWalkCap events.
ORF_HD float beckmann_ms_dielectric_eval(f3 w_i, f3 w_o, float eta, float ax, float ay, uint32_t& rng)
{
    f3 w_r = -w_i;
    float hr = 1.0f + height_inv_cdf(0.9999f);
    float sum = 0.0f;
    bool wr_out = (-w_r.z) > 0.0f;
    const bool wo_out = w_o.z > 0.0f;
    for (int order = 0; order < kMsWalkCap; ++order) {
        const float z = rnd_tea(rng);
        hr = wr_out ? beckmann_sample_height(w_r, hr, z, ax, ay) : -beckmann_sample_height(-w_r, -hr, z, ax, ay);
        if (hr == kRayMax || hr == -kRayMax || __builtin_isinf(hr)) break;
        const float phase = beckmann_eval_dielectric_phase(-w_r, w_o, wr_out, wo_out, eta, ax, ay);
        const float shadowing = wo_out ? ms_masking_height(BECKMANN_DISTRIBUTION, w_o, hr, ax, ay)
                                       : ms_masking_height(BECKMANN_DISTRIBUTION, -w_o, -hr, ax, ay);
        const float I = phase * shadowing;
        if (__builtin_isfinite(I)) sum += I;
        w_r = beckmann_sample_dielectric_phase(-w_r, eta, rng, wr_out, ax, ay);
        if (hr != hr || w_r.z != w_r.z) return 0.0f;
    }
    return sum;
}

// One step of the rough dielectric (CU/rough_transparent_shader.cu:66-225,
// visible-normal, Walter or multiple-scattering model): Beer absorption with Russian roulette when leaving the
// medium, a visible microfacet normal, Fresnel-weighted choice between
// reflection and refraction (get_rough_glass_rays, :39-55), weight G1(w_o).
// Returns false where the reference returns 0; `factor` multiplies the throughput.
ORF_HD bool dielectric_step(f3 d, float t_hit, f3 normal, f3 ffn, const MaterialGPU& m, uint32_t& rng,
                            f3& new_dir, f3& factor)
{
    f3 w_i = -d;
    f3 beam = mk3(1.0f);
    if (dot(-d, normal) < 0.0f) {
        f3 e = (-t_hit) * m.glass_absorption;
        beam = mk3(orf_expf(e.x), orf_expf(e.y), orf_expf(e.z));
        float prob = (beam.x + beam.y + beam.z) / 3.0f;
        if (rnd_tea(rng) >= prob) return false;
        beam = beam / prob;
    }
    float z1 = rnd_tea(rng);
    float z2 = rnd_tea(rng);
    float a = m.roughness.x;
    int dist = m.normal_distribution;
    if (m.microfacet_model == MULTISCATTERING_MODEL) {
        // CU/rough_transparent_shader.cu:98-116: the walk starts from the
        // un-flipped normal's frame (z1, z2 above are drawn and unused, as there)
        f3 lo;
        const bool ok = dist == GGX_DISTRIBUTION
                            ? ggx_ms_dielectric_sample(to_local(w_i, normal), 1.0f / m.ior, a, a, rng, lo)
                            : beckmann_ms_dielectric_sample(to_local(w_i, normal), 1.0f / m.ior, a, a, rng, lo);
        if (!ok) return false;
        new_dir = to_world(lo, normal);
        factor = beam;
        return true;
    }
    const bool walter = m.microfacet_model == WALTER_MODEL;
    f3 mn = walter ? sample_normal_walter(ffn, z1, z2, a, dist) : sample_visible_normal(w_i, ffn, a, a, z1, z2, dist);
    float eta = dot(normal, w_i) < 0.0f ? m.ior : 1.0f / m.ior;
    float g_i = walter ? masking_g1_iso(w_i, mn, ffn, a, dist) : masking_g1(w_i, mn, ffn, a, a, dist);
    if (g_i <= 0.0f) return false;
    float ct = fabsf(dot(mn, -d));
    float st2 = eta * eta * (1.0f - ct * ct);
    float ctt = sqrtf(1.0f - st2);
    float R = st2 < 1.0f ? fresnel_R3(ct, ctt, eta) : 1.0f;
    float rr = rnd_tea(rng);
    f3 dir = rr > R ? eta * d + mn * (eta * ct - ctt) : d - (mn * 2.0f) * dot(mn, d);
    float g_o = walter ? masking_g1_iso(dir, mn, ffn, a, dist) : masking_g1(dir, mn, ffn, a, a, dist);
    if (g_o <= 0.0f) return false;
    new_dir = dir;
    factor = beam * (walter ? walter_weight(w_i, mn, ffn, normal, g_i, g_o) : g_o);
    return true;
}

// Power-heuristic weight of strategy a against b.
ORF_HD float mis_power(float pa, float pb)
{
    float a2 = pa * pa, b2 = pb * pb;
    return a2 + b2 > 0.0f ? a2 / (a2 + b2) : 0.0f;
}

// Shading frame of a hit: interpolated vertex normal (CU/triangle_mesh.cu:69),
// normalised again as rtTransformNormal + normalize do in the closest-hit programs.
ORF_HD f3 shading_normal(const SceneGPU& sc, int prim, float b, float g, f3 e1, f3 e2, int4& sh)
{
    sh = sc.shade[prim];
    f3 n;
    if (sh.x < 0 || sh.y < 0 || sh.z < 0) {
        n = normalize(cross(e1, e2));
    } else {
        float4 a0 = sc.normals[sh.x], a1 = sc.normals[sh.y], a2 = sc.normals[sh.z];
        f3 n0 = mk3(a0.x, a0.y, a0.z), n1 = mk3(a1.x, a1.y, a1.z), n2 = mk3(a2.x, a2.y, a2.z);
        n = normalize(n1 * b + n2 * g + n0 * (1.0f - b - g));
    }
    return normalize(n);
}

}  // namespace orf
