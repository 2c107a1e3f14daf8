// orf_path.h -- the integrator's per-vertex shading step and the camera path
// loop built on it (device; also host-callable).
//
// The recursion of the reference's closest-hit programs (rtTrace from inside
// CU/diffuse_shader.cu, CU/metallic_shader.cu, CU/rough_transparent_shader.cu)
// becomes a loop carrying a throughput; the TEA state is threaded through as the
// reference threads prd.seed, so every draw lands at the same stream position.
// Depth rules are the reference's: Lambert shades while depth <= max_depth
// (CU/diffuse_shader.cu:45), conductor and dielectric while depth < max_depth
// (CU/metallic_shader.cu:47, CU/rough_transparent_shader.cu:68); emitters and the
// environment answer at any depth.  emit_light follows CU/path_tracer.cu:46,
// CU/diffuse_shader.cu:80,96 and CU/metallic_shader.cu:127.
//
// shade_vertex() is the one place the shading arithmetic lives: the megakernel
// calls it inside its loop, the wavefront kernels call it on queued hits, so
// both forms produce bit-identical per-sample radiance.
#pragma once

#include "orf_shade.h"
#include "orf_sss.h"
#include "orf_trace.h"

namespace orf {

struct PathStats {
    unsigned ext = 0, shd = 0;
};

struct PathState {
    f3 o, d;             // current ray
    f3 thr;              // throughput
    f3 L;                // radiance gathered so far
    uint32_t rng;        // TEA state (prd.seed)
    Seed64 s64;          // MD5 state of the translucent programs (prd.seed64)
    int depth;           // prd.depth
    int emit;            // prd.emit_light
    float prev_bsdf_pdf; // > 0: the ray left a Lambert vertex that sampled lights (env MIS)
};

struct ShadowRay {
    f3 o, d;
    float tmax;
    f3 contrib;  // added to L when the segment is unoccluded
    int emit;    // >= 0: the path's emit flag becomes this when unoccluded
    bool valid;
};

// Applies an unoccluded shadow segment to the path.
ORF_HD void shadow_unoccluded(PathState& ps, const ShadowRay& sr)
{
    ps.L += sr.contrib;
    if (sr.emit >= 0) ps.emit = sr.emit;
}

// Camera ray of pixel (x, y) of a w x h film for sample `frame`
// (CU/path_tracer.cu:45-55); returns the seeded TEA state.  `seed0`, when
// given, receives the pixel seed before the jitter draws (both words of the
// path's 64-bit state, :49).
ORF_HD uint32_t camera_ray(const CameraGPU& cam, int w, int h, int x, int y, unsigned frame, f3& dir,
                           uint32_t* seed0 = nullptr)
{
    uint32_t rng = tea<16>((unsigned)w * (unsigned)y + (unsigned)x, frame);
    if (seed0) *seed0 = rng;
    float jx = rnd_tea(rng);
    float jy = rnd_tea(rng);
    float ipx = ((float)x + jx) / (float)w * 2.0f - 1.0f;
    float ipy = ((float)y + jy) / (float)h * 2.0f - 1.0f;
    dir = normalize(ipx * cam.U + ipy * cam.V + cam.W);
    return rng;
}

ORF_HD void path_begin(PathState& ps, f3 eye, f3 dir, uint32_t rng, Seed64 s64)
{
    ps.o = eye;
    ps.d = dir;
    ps.thr = mk3(1.0f);
    ps.L = mk3(0.0f);
    ps.rng = rng;
    ps.s64 = s64;
    ps.depth = 0;
    ps.emit = 1;
    ps.prev_bsdf_pdf = -1.0f;
}

// Material of a hit through the per-render material table (identity when null).
ORF_HD const MaterialGPU& hit_material(const SceneGPU& sc, const int* remap, int id)
{
    return sc.materials[remap ? remap[id] : id];
}

// Environment answer of a ray that left the scene (CU/envmap_background.cu:20-23 + MIS).
ORF_HD void shade_miss(const SceneGPU& sc, PathState& ps)
{
    if (!sc.has_env) return;
    f3 E = env_lookup(sc.env, ps.d);
    if (ps.prev_bsdf_pdf > 0.0f && sc.env.sampled)
        E = E * mis_power(ps.prev_bsdf_pdf, env_pdf(sc.env, ps.d) / (float)sc.nlights);
    ps.L += ps.thr * E;
}

// Geometry of a hit: position, shading normal, the normal facing the ray, and
// the shade record (vertex-normal indices, geometry material id).
struct HitGeom {
    f3 p, normal, ffn;
    int4 sh;
};

template <bool kEmit = true>
ORF_HD HitGeom hit_geometry(const SceneGPU& sc, const PathState& ps, const Hit& h)
{
    HitGeom g;
    g.p = ps.o + h.t * ps.d;
    if (kEmit && h.leaf < 0) {  // disk or sphere emitter
        const int i = -1 - h.leaf;
        g.sh = sc.shade[h.prim];
        g.normal = analytic_normal(sc.analytic[2 * i], sc.analytic[2 * i + 1], ps.o, ps.d, h.t);
        g.ffn = faceforward(g.normal, -ps.d, g.normal);
        return g;
    }
    const float4* tg = sc.tri + 3 * (size_t)h.leaf;
    float4 q1 = tg[1], q2 = tg[2];
    g.normal = shading_normal(sc, h.prim, h.b, h.g, mk3(q1.x, q1.y, q1.z), mk3(q2.x, q2.y, q2.z), g.sh);
    g.ffn = faceforward(g.normal, -ps.d, g.normal);
    return g;
}

template <unsigned kFeat = kFeatAll>
ORF_HD bool shade_with(const SceneGPU& sc, const MaterialGPU& mat, PathState& ps, const Hit& h, const HitGeom& g,
                       ShadowRay& sr, unsigned* stk = nullptr, int stride = 1, SssGrad* sg = nullptr);

// Shade one surface vertex.  Returns true when the path continues with ps.o/ps.d
// as its next ray; `sr` receives the NEE shadow segment, if any.  kFeat: the
// material code compiled in (kFeat* bits); the caller guarantees it covers the
// scene's materials.  (stk, stride): a traversal stack the shading may use for
// rays of its own (the translucent gather), or null.
template <unsigned kFeat = kFeatAll>
ORF_HD bool shade_vertex(const SceneGPU& sc, const int* remap, PathState& ps, const Hit& h, ShadowRay& sr,
                         unsigned* stk = nullptr, int stride = 1)
{
    HitGeom g = hit_geometry<kFeat == kFeatAll>(sc, ps, h);  // emitters only in scenes of the all-feature variant
    return shade_with<kFeat>(sc, hit_material(sc, remap, g.sh.w), ps, h, g, sr, stk, stride);
}

// Shade a vertex with material `mat` (the material programs' closest hits).
// `sg` (translucent materials): receives what the vertex contributes to the
// derivatives with respect to the material's absorption and scattering.
template <unsigned kFeat>
ORF_HD bool shade_with(const SceneGPU& sc, const MaterialGPU& mat, PathState& ps, const Hit& h, const HitGeom& g,
                       ShadowRay& sr, unsigned* stk, int stride, SssGrad* sg)
{
    constexpr bool kDiel = (kFeat & kFeatDielectric) != 0, kSss = (kFeat & kFeatTranslucent) != 0;
    sr.valid = false;
    sr.emit = -1;
    const float eps = sc.scene_epsilon;
    f3 d = ps.d;
    f3 p = g.p, normal = g.normal, ffn = g.ffn;
    if (mat.kind == MAT_EMITTER) {
        if (dot(normal, -d) > 0.0f && ps.emit) ps.L += ps.thr * mat.radiance;
        return false;
    }
    if (mat.kind == MAT_LAMBERT) {
        if (ps.depth > sc.max_depth) return false;
        f3 rho = mat.diffuse_color;
        unsigned li = (unsigned)((float)sc.nlights * rnd_tea(ps.rng));
        if (sc.nlights > 0) {
            if (li >= (unsigned)sc.nlights) li = sc.nlights - 1;
            float light_pdf = 1.0f / (float)sc.nlights;
            LightGPU lt = sc.lights[li];
            f3 wl = mk3(0.0f), rad = mk3(0.0f);
            float dist = 0.0f, penv;
            sample_light<kFeat == kFeatAll>(sc, lt, p, ps.rng, wl, rad, dist, penv);
            rad = rad / light_pdf;
            float cs = dot(ffn, wl);
            if (cs > 0.0f) {
                f3 c = rad * kInvPi * cs * rho;
                if (penv >= 0.0f) c = c * mis_power(penv * light_pdf, cs * kInvPi);
                sr.o = p;
                sr.d = wl;
                sr.tmax = dist - eps;
                sr.contrib = ps.thr * c;
                sr.valid = true;
                ps.emit = 0;
            }
        }
        float prob = (rho.x + rho.y + rho.z) / 3.0f;
        float xi = rnd_tea(ps.rng);
        if (!(xi < prob)) return false;
        f3 nd = sample_cosine(ffn, ps.rng);
        ps.prev_bsdf_pdf = sc.env.sampled ? dot(ffn, nd) * kInvPi : -1.0f;
        ps.thr = ps.thr * (rho / prob);
        ps.o = p;
        ps.d = nd;
        ps.depth++;
        return true;
    }
    if (mat.kind == MAT_CONDUCTOR) {
        if (ps.depth >= sc.max_depth) return false;
        f3 wo, w;
        if (kDiel && mat.microfacet_model == MULTISCATTERING_MODEL) {
            // CU/metallic_shader.cu:66-82: the microsurface walk (GGX or Beckmann) in
            // the frame of the unflipped normal; its weight multiplies the continuation
            f3 lo;
            const bool ok = mat.normal_distribution == GGX_DISTRIBUTION
                                ? ggx_ms_conductor_sample(to_local(-d, normal), mat.eta, mat.roughness.x,
                                                          mat.roughness.y, ps.rng, lo, w)
                                : beckmann_ms_conductor_sample(to_local(-d, normal), mat.eta, mat.roughness.x,
                                                               mat.roughness.y, ps.rng, lo, w);
            if (!ok) return false;
            wo = to_world(lo, normal);
        } else {
            float z1 = rnd_tea(ps.rng);
            float z2 = rnd_tea(ps.rng);
            conductor_step(-d, ffn, mat, z1, z2, wo, w);
            if (w.x == 0.0f && w.y == 0.0f && w.z == 0.0f) return false;
        }
        ps.thr = ps.thr * w;
        ps.o = p;
        ps.d = wo;
        ps.depth++;
        ps.emit = 1;
        ps.prev_bsdf_pdf = -1.0f;
        return true;
    }
    if (kDiel && mat.kind == MAT_DIELECTRIC) {
        if (ps.depth >= sc.max_depth) return false;
        f3 nd, f;
        if (!dielectric_step(d, h.t, normal, ffn, mat, ps.rng, nd, f)) return false;
        ps.thr = ps.thr * f;
        ps.o = p;
        ps.d = nd;
        ps.depth++;
        ps.emit = 1;
        ps.prev_bsdf_pdf = -1.0f;
        return true;
    }
    if (kDiel && mat.kind == MAT_TRANSPARENT) {
        // CU/transparent_shader.cu:51-114 (smooth glass, no diffuse part): depth
        // test depth > max_depth, Beer absorption with Russian roulette inside,
        // Fresnel choice between the mirror and the refracted direction.
        if (ps.depth > sc.max_depth) return false;
        f3 beam = mk3(1.0f);
        float eta = 1.0f / mat.ior;
        f3 n = normal;
        float c = dot(-d, n);
        if (c < 0.0f) {
            f3 e = (-h.t) * mat.glass_absorption;
            beam = mk3(orf_expf(e.x), orf_expf(e.y), orf_expf(e.z));
            float prob = (beam.x + beam.y + beam.z) / 3.0f;
            if (rnd_tea(ps.rng) >= prob) return false;
            beam = beam / prob;
            eta = mat.ior;
            n = -n;
            c = -c;
        }
        float R = 1.0f, ct = 0.0f;
        float st2 = eta * eta * (1.0f - c * c);
        if (st2 < 1.0f) {
            ct = sqrtf(1.0f - st2);
            R = fresnel_R3(c, ct, eta);
        }
        float xi = rnd_tea(ps.rng);
        ps.d = xi < R ? d - (n * 2.0f) * dot(n, d) : eta * d + n * (eta * c - ct);
        ps.thr = ps.thr * beam;
        ps.o = p;
        ps.depth++;
        ps.emit = 1;
        ps.prev_bsdf_pdf = -1.0f;
        return true;
    }
    if (kDiel && mat.kind == MAT_NORMAL) {
        // CU/normal_shader.cu:25-29: the shading normal mapped to [0, 1]^3; the path ends
        ps.L += ps.thr * (normal * 0.5f + 0.5f);
        return false;
    }
    if (kDiel && mat.kind == MAT_ROUGH_DIFFUSE) {
        // CU/rough_diffuse_shader.cu:40-290: a Lambert base under a rough
        // dielectric interface.  NEE adds the transmitted diffuse and the
        // microfacet-reflected terms (shadow segment to the full light distance,
        // emission off once a light is seen); the continuation is the diffuse
        // bounce about the sampled microfacet normal with probability 1 - R_i,
        // else the mirror about it, weighted as the reference weights them (not
        // divided by the choice probability).  The multiple-scattering model
        // replaces both NEE terms by walk estimates (diffuse and interface
        // reflection, drawn before the shadow test as the rough translucent
        // boundary's are) and both continuations by walks: the diffuse walk
        // (weight rho^k n12^2 T / (1 - rho r10), T at its last visible normal)
        // and the reflection walk with the interface as a conductor of index
        // n12, weight 1, in the un-flipped normal's frame (:252-268).
        if (ps.depth > sc.max_depth) return false;
        const f3 w_i = -d;
        const float n12 = 1.0f / mat.ior;
        const float r10 = two_C1(mat.ior);
        const float ax = mat.roughness.x, ay = mat.roughness.y;
        const int nd = mat.normal_distribution;
        const f3 rho = mat.diffuse_color;
        const f3 denom = mk3(1.0f - rho.x * r10, 1.0f - rho.y * r10, 1.0f - rho.z * r10);
        const bool ms = mat.microfacet_model == MULTISCATTERING_MODEL;
        const float z1 = rnd_tea(ps.rng);
        const float z2 = rnd_tea(ps.rng);
        const f3 mn = sample_visible_normal(w_i, ffn, ax, ay, z1, z2, nd);
        const float ci = dot(w_i, mn);
        float Ri = 1.0f;
        float st2 = n12 * n12 * (1.0f - ci * ci);
        if (st2 < 1.0f) Ri = fresnel_R3(ci, sqrtf(1.0f - st2), n12);
        const float Ti = 1.0f - Ri;
        unsigned li = (unsigned)((float)sc.nlights * rnd_tea(ps.rng));
        if (sc.nlights > 0) {
            if (li >= (unsigned)sc.nlights) li = sc.nlights - 1;
            const float light_pdf = 1.0f / (float)sc.nlights;
            f3 wl = mk3(0.0f), rad = mk3(0.0f);
            float dist = 0.0f, penv;
            sample_light<kFeat == kFeatAll>(sc, sc.lights[li], p, ps.rng, wl, rad, dist, penv);
            rad = rad / light_pdf;
            const float cl = dot(ffn, wl);
            // the environment reaches this material through escaping rays only
            if (cl > 0.0f && penv < 0.0f) {
                f3 c;
                if (ms) {
                    // :126-139
                    const f3 lwi = to_local(w_i, ffn), lwl = to_local(wl, ffn);
                    const f3 fd = ms_diffuse_eval(nd, lwi, lwl, ax, ay, rho, ps.rng);
                    c = rad * (fd + ms_conductor_eval(nd, lwi, lwl, mat.eta, ax, ay, ps.rng));
                } else {
                    const f3 rn = normalize(w_i + wl);
                    const float cr = dot(rn, wl);
                    float Rr = 1.0f;
                    st2 = n12 * n12 * (1.0f - cr * cr);
                    if (st2 < 1.0f) Rr = fresnel_R3(cr, sqrtf(1.0f - st2), n12);
                    const float cd = dot(mn, wl);
                    float Rd = 1.0f;
                    st2 = n12 * n12 * (1.0f - cd * cd);
                    if (st2 < 1.0f) Rd = fresnel_R3(cd, sqrtf(1.0f - st2), n12);
                    const float Td = 1.0f - Rd;
                    const float gi = masking_g1(w_i, rn, ffn, ax, ay, nd);
                    const float go = masking_g1(wl, rn, ffn, ax, ay, nd);
                    const float god = masking_g1(wl, mn, ffn, ax, ay, nd);
                    const float D = eval_visible_normal(w_i, rn, ffn, ax, ay, nd);
                    const float wr = gi * go * D / (4.0f * fabsf(dot(w_i, ffn) * dot(wl, ffn)));
                    const f3 diff = rad * kInvPi * cl * rho * n12 * n12 * Ti * Td / denom * god;
                    c = diff + rad * Rr * wr;
                }
                sr.o = p;
                sr.d = wl;
                sr.tmax = dist;  // the reference's shadow ray runs to the light itself (:96)
                sr.contrib = ps.thr * c;
                sr.emit = 0;
                sr.valid = true;
            }
        }
        const float xi = rnd_tea(ps.rng);
        f3 nd_dir, w;
        if (ms) {
            f3 lo;
            if (xi > Ri) {
                // :187-216
                f3 lm, k;
                ms_diffuse_sample(nd, to_local(w_i, ffn), ax, ay, rho, ps.rng, lo, lm, k);
                nd_dir = to_world(lo, ffn);
                const float cr = dot(nd_dir, to_world(lm, ffn));
                float Rr = 1.0f;
                st2 = n12 * n12 * (1.0f - cr * cr);
                if (st2 < 1.0f) Rr = fresnel_R3(cr, sqrtf(1.0f - st2), n12);
                w = k * (n12 * n12 * (1.0f - Rr)) / denom;
            } else {
                // :252-268
                f3 k;
                const bool ok = nd == GGX_DISTRIBUTION
                                    ? ggx_ms_conductor_sample(to_local(w_i, normal), mat.eta, ax, ay, ps.rng, lo, k)
                                    : beckmann_ms_conductor_sample(to_local(w_i, normal), mat.eta, ax, ay, ps.rng, lo, k);
                if (!ok) return false;
                nd_dir = to_world(lo, normal);
                w = mk3(1.0f);
            }
        } else if (xi > Ri) {
            nd_dir = sample_cosine(mn, ps.rng);
            const float cr = dot(nd_dir, mn);
            const float gom = masking_g1(nd_dir, mn, ffn, ax, ay, nd);
            if (gom < 0.0f) return false;
            float Rr = 1.0f;
            st2 = n12 * n12 * (1.0f - cr * cr);
            if (st2 < 1.0f) Rr = fresnel_R3(cr, sqrtf(1.0f - st2), n12);
            w = rho * n12 * n12 * (1.0f - Rr) / denom * gom;
        } else {
            nd_dir = d - (mn * 2.0f) * dot(mn, d);
            const float gom = masking_g1(nd_dir, mn, ffn, ax, ay, nd);
            if (gom < 0.0f) return false;
            w = mk3(gom);
        }
        ps.thr = ps.thr * w;
        ps.o = p;
        ps.d = nd_dir;
        ps.depth++;
        ps.prev_bsdf_pdf = -1.0f;
        return true;
    }
    if (kDiel && mat.kind == MAT_FLAT) {
        // CU/flat_shader.cu:40-127: the flat colour; with flat_shadow it is
        // multiplied by the light's visibility V, and with highlight a visible
        // light whose microfacet reflection (luminance of Le G1(i) G1(o) D_wi R /
        // (4 |i.n l.n|)) exceeds the threshold shows the highlight colour.  The
        // colour for V = 1 is A, for V = 0 it is B; B is added now and A - B when
        // the shadow segment resolves unoccluded.
        if (ps.depth > sc.max_depth) return false;
        const f3 w_i = -d;
        rnd_tea(ps.rng);  // z1, z2 of a microfacet normal the shader samples and never uses
        rnd_tea(ps.rng);
        const f3 B = (mat.flags & kFlatShadow) ? mk3(0.0f) : mat.diffuse_color;
        unsigned li = (unsigned)((float)sc.nlights * rnd_tea(ps.rng));
        if (sc.nlights > 0) {
            if (li >= (unsigned)sc.nlights) li = sc.nlights - 1;
            const float light_pdf = 1.0f / (float)sc.nlights;
            f3 wl = mk3(0.0f), rad = mk3(0.0f);
            float dist = 0.0f, penv;
            sample_light<kFeat == kFeatAll>(sc, sc.lights[li], p, ps.rng, wl, rad, dist, penv);
            rad = rad / light_pdf;
            if (dot(ffn, wl) > 0.0f) {
                f3 A = mat.diffuse_color;
                if (mat.flags & kFlatHighlight) {
                    const f3 rn = normalize(w_i + wl);
                    const float cr = dot(rn, wl), eta = 1.0f / mat.ior;
                    float R = 1.0f;
                    const float st2 = eta * eta * (1.0f - cr * cr);
                    if (st2 < 1.0f) R = fresnel_R3(cr, sqrtf(1.0f - st2), eta);
                    const float ax = mat.roughness.x, ay = mat.roughness.y;
                    const int nd = mat.normal_distribution;
                    const float gi = masking_g1(w_i, rn, ffn, ax, ay, nd);
                    const float go = masking_g1(wl, rn, ffn, ax, ay, nd);
                    const float D = eval_visible_normal(w_i, rn, ffn, ax, ay, nd);
                    const float wr = gi * go * D * R / (4.0f * fabsf(dot(w_i, ffn) * dot(wl, ffn)));
                    const f3 spec = rad * wr;
                    const float lum = 0.2126f * spec.x + 0.7152f * spec.y + 0.0722f * spec.z;
                    if (lum > mat.threshold) A = mat.radiance;
                }
                ps.L += ps.thr * B;
                sr.o = p;
                sr.d = wl;
                sr.tmax = dist - eps;
                sr.contrib = ps.thr * (A - B);
                sr.valid = true;
                return false;
            }
        }
        ps.L += ps.thr * B;
        return false;
    }
    if (kDiel && mat.kind == MAT_INTERFACE) {
        // Lambertian interface (CU/lambertian_interface_shader.cu:51-173): a Lambert
        // base under a smooth dielectric of index ior.  NEE through both Fresnel
        // transmissions; the continuation is the specular reflection with the
        // probability R_i of the incident Fresnel, else a cosine diffuse bounce.
        if (ps.depth > sc.max_depth) return false;
        f3 rho = mat.diffuse_color;
        float n12 = 1.0f / mat.ior;
        float r10 = two_C1(mat.ior);
        f3 denom = mk3(1.0f - rho.x * r10, 1.0f - rho.y * r10, 1.0f - rho.z * r10);
        float ci = dot(-d, ffn);
        float Ri = fresnel_R_cos(ci, n12);
        float Ti = 1.0f - Ri;
        unsigned li = (unsigned)((float)sc.nlights * rnd_tea(ps.rng));
        if (sc.nlights > 0) {
            if (li >= (unsigned)sc.nlights) li = sc.nlights - 1;
            float light_pdf = 1.0f / (float)sc.nlights;
            LightGPU lt = sc.lights[li];
            f3 wl = mk3(0.0f), rad = mk3(0.0f);
            float dist = 0.0f, penv;
            sample_light<kFeat == kFeatAll>(sc, lt, p, ps.rng, wl, rad, dist, penv);
            rad = rad / light_pdf;
            float cl = dot(ffn, wl);
            if (cl > 0.0f) {
                float Tl = 1.0f - fresnel_R_cos(cl, n12);
                f3 c = rad * kInvPi * cl * rho * n12 * n12 * Tl * Ti / denom;
                if (penv >= 0.0f) c = c * mis_power(penv * light_pdf, cl * kInvPi);
                sr.o = p;
                sr.d = wl;
                sr.tmax = dist;  // CU/lambertian_interface_shader.cu:100 (no epsilon)
                sr.contrib = ps.thr * c;
                sr.valid = true;
                ps.emit = 0;
            }
        }
        float xi = rnd_tea(ps.rng);
        if (xi > Ri) {
            f3 nd = sample_cosine(ffn, ps.rng);
            float cr = dot(nd, ffn);
            float Tr = 1.0f - fresnel_R_cos(cr, n12);
            ps.prev_bsdf_pdf = sc.env.sampled ? cr * kInvPi : -1.0f;
            ps.thr = ps.thr * (rho * n12 * n12 * Tr / denom);
            ps.o = p;
            ps.d = nd;
            ps.depth++;
            return true;
        }
        ps.o = p;
        ps.d = d - (normal * 2.0f) * dot(normal, d);  // reflect about the shading normal
        ps.depth++;
        ps.emit = 1;
        ps.prev_bsdf_pdf = -1.0f;
        return true;
    }
    if (kSss && mat.kind == MAT_TRANSLUCENT) {
        // CU/subsurface_scattering_shader.cu:55-198: Fresnel choice between the
        // specular reflection and the refraction; entering from outside also adds
        // the diffusion term (sss_gather, before the path continues); inside, Beer
        // extinction with Russian roulette first.
        if (ps.depth > sc.max_depth) return false;
        // Every draw here is from the 64-bit stream (RND_64, :8).
        const ScatteringGPU& P = sc.sss[mat.sss];
        f3 no = normal, wo = -d;
        float recip = 1.0f / P.relative_ior;
        float reflect_xi = rnd_accurate(ps.s64);
        float co = dot(wo, no);
        bool inside = co < 0.0f;
        f3 beam = mk3(1.0f);
        if (inside) {
            f3 ext = P.dipole_model == 1 ? P.deltaEddExtinction : P.extinction;
            beam = exp3((-h.t) * ext);
            float prob = (beam.x + beam.y + beam.z) / 3.0f;
            if (rnd_accurate(ps.s64) >= prob) return false;
            beam = beam / prob;
            recip = P.relative_ior;
            no = -no;
            co = -co;
            if (sg) sss_beam_grad(*sg, P, h.t);
        }
        float st2 = recip * recip * (1.0f - co * co);
        float ct = 1.0f, R = 1.0f;
        if (st2 < 1.0f) {
            ct = sqrtf(1.0f - st2);
            R = fresnel_R3(co, ct, recip);
        }
        ps.thr = ps.thr * beam;
        f3 nd;
        if (reflect_xi >= R) {
            if (!inside && mat.sss == sc.probes.sss) {
                ps.L += ps.thr * sss_gather(sc, P, mat.sss, p, no, ps.s64, stk, stride, sg);
                if (sg) {
                    sg->dL_a = ps.thr * sg->dL_a;
                    sg->dL_s = ps.thr * sg->dL_s;
                }
            }
            nd = recip * (co * no - wo) - no * ct;
        } else {
            nd = 2.0f * co * no - wo;
        }
        ps.o = p;
        ps.d = nd;
        ps.depth++;
        ps.emit = 1;
        ps.prev_bsdf_pdf = -1.0f;
        return true;
    }
    if (kSss && mat.kind == MAT_ROUGH_TRANSLUCENT) {
        // CU/rough_subsurface_scattering_shader.cu:64-455: a microfacet boundary
        // (Walter, visible-normal or multiple-scattering model) over the
        // diffusion term.  NEE from outside through the boundary's reflection
        // (the reference's weight G_i G_o D_wi / (4 |i.n| |l.n|) times Fresnel,
        // or a multiple-scattering BSDF estimate); the continuation is the
        // Fresnel choice between refraction (gathering the diffusion term when
        // entering) and reflection; inside, Beer extinction with roulette.
        if (ps.depth > sc.max_depth) return false;
        const ScatteringGPU& P = sc.sss[mat.sss];
        const f3 w_i = -d;
        const float ax = mat.roughness.x, ay = mat.roughness.y;
        const int dist = mat.normal_distribution, model = mat.microfacet_model;
        // reflect_xi, the facet's z1 and z2, the light pick, the extinction
        // roulette and the gather draw from the 64-bit stream (RND_64, :8,
        // :81-84, :108-109, :186-187); the light sample and the multiple-
        // scattering walk and evaluation from the TEA stream
        float recip = 1.0f / P.relative_ior;
        const float reflect_xi = rnd_accurate(ps.s64);
        const float z1 = rnd_accurate(ps.s64), z2 = rnd_accurate(ps.s64);
        // the reference passes (a_x, z1, z2) to microfacet_sample_normal's
        // (z1, z2, alpha) parameters (:96); kept as written
        f3 mn = model == WALTER_MODEL ? sample_normal_walter(ffn, ax, z1, z2, dist)
                                      : sample_visible_normal(w_i, ffn, ax, ay, z1, z2, dist);
        const float cos_i = dot(w_i, mn);
        const bool inside = dot(w_i, normal) < 0.0f;
        // NEE over the reference's lights (the env, a light here, is not one there)
        const int nreal = sc.nlights - (sc.env.sampled ? 1 : 0);
        unsigned li = (unsigned)((float)nreal * rnd_accurate(ps.s64));
        bool nee_reflect = false;
        if (nreal > 0) {
            if (li >= (unsigned)nreal) li = nreal - 1;
            const float light_pdf = 1.0f / (float)nreal;
            f3 wl = mk3(0.0f), rad = mk3(0.0f);
            float ldist = 0.0f, penv;
            sample_light<kFeat == kFeatAll>(sc, sc.lights[li], p, ps.rng, wl, rad, ldist, penv);
            rad = rad / light_pdf;
            const float cl = dot(ffn, wl);
            if (!inside && cl > 0.0f) {
                f3 c;
                if (model == MULTISCATTERING_MODEL) {
                    const f3 lwi = to_local(w_i, ffn), lwl = to_local(wl, ffn);
                    float f = dist == GGX_DISTRIBUTION ? ggx_ms_dielectric_eval(lwi, lwl, recip, ax, ay, ps.rng)
                                                       : beckmann_ms_dielectric_eval(lwi, lwl, recip, ax, ay, ps.rng);
                    c = rad * f;
                } else {
                    f3 rn = normalize(w_i + wl);
                    float cr = dot(rn, wl);
                    float Rr = 1.0f;
                    float st2 = recip * recip * (1.0f - cr * cr);
                    if (st2 < 1.0f) Rr = fresnel_R3(cr, sqrtf(1.0f - st2), recip);
                    float gi = masking_g1(w_i, rn, ffn, ax, ay, dist);
                    float go = masking_g1(wl, rn, ffn, ax, ay, dist);
                    float D = eval_visible_normal(w_i, rn, ffn, ax, ay, dist);
                    float wr = gi * go * D / (4.0f * fabsf(dot(w_i, ffn) * dot(wl, ffn)));
                    c = rad * Rr * wr;
                    nee_reflect = true;
                }
                sr.o = p;
                sr.d = wl;
                sr.tmax = ldist;  // :130 (no epsilon)
                sr.contrib = ps.thr * c;
                sr.valid = true;
            }
        }
        f3 beam = mk3(1.0f);
        f3 nd, weight;
        bool refract;
        if (model == MULTISCATTERING_MODEL) {
            // :360-455: the walk from the un-flipped normal's frame, weight 1
            f3 lo;
            const f3 lwi = to_local(w_i, normal);
            const float eta = 1.0f / P.relative_ior;
            if (!(dist == GGX_DISTRIBUTION ? ggx_ms_dielectric_sample(lwi, eta, ax, ay, ps.rng, lo)
                                           : beckmann_ms_dielectric_sample(lwi, eta, ax, ay, ps.rng, lo)))
                return false;
            nd = to_world(lo, normal);
            refract = dot(nd, ffn) < 0.0f;
            weight = mk3(1.0f);
        }
        if (inside) {
            f3 ext = P.dipole_model == 1 ? P.deltaEddExtinction : P.extinction;
            beam = exp3((-h.t) * ext);
            float prob = (beam.x + beam.y + beam.z) / 3.0f;
            if (rnd_accurate(ps.s64) >= prob) return false;
            beam = beam / prob;
            recip = P.relative_ior;
            if (sg) sss_beam_grad(*sg, P, h.t);
        }
        const f3 no = inside ? -normal : normal;
        if (model != MULTISCATTERING_MODEL) {
            float st2 = recip * recip * (1.0f - cos_i * cos_i);
            float cos_t = 1.0f, R = 1.0f;
            if (st2 < 1.0f) {
                cos_t = sqrtf(1.0f - st2);
                R = fresnel_R3(cos_i, cos_t, recip);
            }
            refract = reflect_xi >= R;
            nd = refract ? recip * (cos_i * mn - w_i) - mn * cos_t : 2.0f * cos_i * mn - w_i;
            float w;
            if (model == WALTER_MODEL) {
                float gi = masking_g1_iso(w_i, mn, ffn, ax, dist);
                float go = masking_g1_iso(nd, mn, ffn, ax, dist);
                w = fabsf(dot(w_i, mn)) * gi * go / (fabsf(dot(w_i, ffn)) * fabsf(dot(ffn, mn)));
            } else {
                w = masking_g1(nd, mn, ffn, ax, ay, dist);
            }
            weight = mk3(w);
        }
        ps.thr = ps.thr * beam;
        if (refract) {
            if (!inside && mat.sss == sc.probes.sss) {
                ps.L += ps.thr * weight * sss_gather(sc, P, mat.sss, p, no, ps.s64, stk, stride, sg);
                if (sg) {
                    sg->dL_a = ps.thr * weight * sg->dL_a;
                    sg->dL_s = ps.thr * weight * sg->dL_s;
                }
            }
            ps.emit = 1;
        } else if (model == MULTISCATTERING_MODEL) {
            ps.emit = 1;
        } else if (nee_reflect) {
            sr.emit = 1;  // :165: a visible light sample re-arms emission on the reflection
        }
        ps.thr = ps.thr * weight;
        ps.o = p;
        ps.d = nd;
        ps.depth++;
        ps.prev_bsdf_pdf = -1.0f;
        return true;
    }
    return false;  // materials without a radiance program here
}

// The whole path of one camera sample (megakernel form).
template <unsigned kFeat = kFeatAll>
ORF_HD f3 path_radiance(const SceneGPU& sc, const int* remap, f3 o, f3 d, uint32_t rng, Seed64 s64, unsigned* stk,
                        int stride, PathStats& st)
{
    PathState ps;
    path_begin(ps, o, d, rng, s64);
    for (;;) {
        Hit h;
        st.ext++;
        if (!trace<false, kFeat == kFeatAll, (kFeat != 0u)>(sc, ps.o, ps.d, sc.scene_epsilon, kRayMax, h, stk,
                                                            stride)) {
            shade_miss(sc, ps);
            break;
        }
        ShadowRay sr;
        bool more = shade_vertex<kFeat>(sc, remap, ps, h, sr, stk, stride);
        if (sr.valid) {
            Hit hs;
            st.shd++;
            if (!trace<true, kFeat == kFeatAll, (kFeat != 0u)>(sc, sr.o, sr.d, sc.scene_epsilon, sr.tmax, hs, stk,
                                                               stride))
                shadow_unoccluded(ps, sr);
        }
        if (!more) break;
    }
    return ps.L;
}

}  // namespace orf
