// orf_probe_sample.h -- one incident-light sample of a translucent object's probe
// (device; also host-callable).
//
// As CU/sample_camera.cu:109-212 does for its per-frame probes: with probability
// 1/2 a light-source sample (uniform light pick, shadow ray to the light), else a
// cosine-sampled path whose first segment counts only light that has already
// bounced (direct light is the other strategy's); Le is divided by 1/2.  The
// probe keeps Le, the refracted incident direction w12 and T12 = 1 - R (times
// the microfacet boundary's transmission weight for a rough translucent
// material, whose probe normal is then a sampled microfacet normal).  Random
// streams as the reference's: t = tea<16>(probe, frame) and the 64-bit state
// (t, t) (:54-57); the three 64-bit draws that place the reference's probe
// (triangle pick and barycentrics, :58,77-78) are drawn and unused, the probe
// positions being fixed here, so the later 64-bit draws sit where the
// reference's do; the strategy choice is the LCG draw rnd(t) (:116), the light
// pick and the facet draws are 64-bit (:122, :135-136, :178-179), the light
// sample and the cosine direction TEA (:124, :161).
#pragma once

#include "orf_path.h"

namespace orf {

// Microfacet normal of a rough boundary for a probe lit from wi (as
// CU/sample_camera.cu:130-151 and :173-191 draw it; the Walter branch keeps the
// reference's (a_x, z1, z2) argument order).
ORF_HD f3 probe_facet(const MaterialGPU& M, f3 wi, f3 n, Seed64& t64)
{
    float z1 = rnd_accurate(t64);
    float z2 = rnd_accurate(t64);
    if (M.microfacet_model == WALTER_MODEL) return sample_normal_walter(n, M.roughness.x, z1, z2, M.normal_distribution);
    return sample_visible_normal(wi, n, M.roughness.x, M.roughness.y, z1, z2, M.normal_distribution);
}

// Outputs: L_out = Le; w12_out = (refracted direction, T12 x the rough boundary's
// weight); sn_out = the normal the light entered through.
ORF_HD void probe_sample(const SceneGPU& sc, const ScatteringGPU& P, const MaterialGPU& M, int j, unsigned frame,
                         unsigned* stk, int stride, float4& L_out, float4& w12_out, float4& sn_out)
{
    const ProbesGPU& pr = sc.probes;
    const bool rough = M.kind == MAT_ROUGH_TRANSLUCENT;
    float4 pp = pr.pos[j], nn = pr.nrm[j];
    f3 pos = mk3(pp.x, pp.y, pp.z), n = mk3(nn.x, nn.y, nn.z);
    uint32_t t = tea<16>((unsigned)j, frame);
    Seed64 t64 = seed64(t);
    for (int k = 0; k < 3; ++k) rnd_accurate(t64);  // the reference's probe placement
    float recip = 1.0f / P.relative_ior;
    f3 Le = mk3(0.0f), wi = n, sn = n;
    if (rnd_lcg(t) < 0.5f) {
        unsigned li = (unsigned)((float)sc.nlights * rnd_accurate(t64));
        if (sc.nlights > 0) {
            if (li >= (unsigned)sc.nlights) li = sc.nlights - 1;
            float light_pdf = 1.0f / (float)sc.nlights;
            f3 wl = mk3(0.0f), rad = mk3(0.0f);
            float dist = 0.0f, penv;
            sample_light(sc, sc.lights[li], pos, t, wl, rad, dist, penv);
            wi = wl;
            if (rough) sn = probe_facet(M, wi, n, t64);
            float cs = fmaxf(dot(wl, n), 0.0f);
            if (cs > 0.0f) {
                Hit hs;
                if (!trace<true>(sc, pos, wl, sc.scene_epsilon, dist, hs, stk, stride)) Le = rad * cs / light_pdf;
            }
        } else if (rough) {
            sn = probe_facet(M, wi, n, t64);
        }
    } else {
        wi = sample_cosine(n, t);
        if (rough) sn = probe_facet(M, wi, n, t64);
        PathState ps;
        path_begin(ps, pos, wi, t, t64);
        ps.emit = 0;  // an emitter seen directly is the light-source strategy's
        for (bool first = true;; first = false) {
            Hit h;
            if (!trace<false>(sc, ps.o, ps.d, sc.scene_epsilon, kRayMax, h, stk, stride)) {
                if (!first || !sc.env.sampled) shade_miss(sc, ps);
                break;
            }
            ShadowRay sr;
            bool more = shade_vertex(sc, nullptr, ps, h, sr);
            if (sr.valid) {
                Hit hs;
                if (!trace<true>(sc, sr.o, sr.d, sc.scene_epsilon, sr.tmax, hs, stk, stride)) shadow_unoccluded(ps, sr);
            }
            if (!more) break;
        }
        Le = ps.L * kPi;
    }
    Le = Le / 0.5f;
    float ci = fmaxf(dot(wi, sn), 0.0f);
    float st2 = recip * recip * (1.0f - ci * ci);
    float ct = sqrtf(1.0f - st2);
    f3 w12 = recip * (ci * sn - wi) - sn * ct;
    float T12 = 1.0f - fresnel_R3(ci, ct, recip);
    float weight = T12;
    if (rough && M.microfacet_model != MULTISCATTERING_MODEL) {
        // :240-254: the boundary's transmission weight
        const float ax = M.roughness.x, ay = M.roughness.y;
        const int dist = M.normal_distribution;
        float go = masking_g1(w12, sn, n, ax, ay, dist);
        float w = go;
        if (M.microfacet_model == WALTER_MODEL) {
            float gi = masking_g1(wi, sn, n, ax, ay, dist);
            w = fabsf(dot(wi, sn)) * gi / (fabsf(dot(wi, n)) * fabsf(dot(n, sn))) * go;
        }
        weight = w * T12;
    }
    L_out = make_float4(Le.x, Le.y, Le.z, 0.0f);
    w12_out = make_float4(w12.x, w12.y, w12.z, weight);
    sn_out = make_float4(sn.x, sn.y, sn.z, 0.0f);
}

}  // namespace orf
