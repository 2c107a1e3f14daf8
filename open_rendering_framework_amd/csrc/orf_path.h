// orf_path.h -- one camera path of the integrator pass (device; also host-callable).
//
// The recursion of the reference's closest-hit programs (rtTrace from inside
// CU/diffuse_shader.cu, CU/metallic_shader.cu, CU/rough_transparent_shader.cu)
// becomes a loop carrying a throughput; the TEA state is threaded through as the
// reference threads prd.seed, so every draw lands at the same stream position.
// Depth rules are the reference's: Lambert shades while depth <= max_depth
// (CU/diffuse_shader.cu:125), conductor and dielectric while depth < max_depth
// (CU/metallic_shader.cu:47, CU/rough_transparent_shader.cu:68); emitters and the
// environment answer at any depth.  emit_light follows CU/path_tracer.cu:46,
// CU/diffuse_shader.cu:160,176 and CU/metallic_shader.cu:127.
#pragma once

#include "orf_shade.h"
#include "orf_trace.h"

namespace orf {

struct PathStats {
    unsigned ext = 0, shd = 0;
};

// Camera ray of pixel (x, y) of a w x h film for sample `frame`
// (CU/path_tracer.cu:45-55); returns the seeded TEA state.
ORF_HD uint32_t camera_ray(const CameraGPU& cam, int w, int h, int x, int y, unsigned frame, f3& dir)
{
    uint32_t rng = tea<16>((unsigned)w * (unsigned)y + (unsigned)x, frame);
    float jx = rnd_tea(rng);
    float jy = rnd_tea(rng);
    float ipx = ((float)x + jx) / (float)w * 2.0f - 1.0f;
    float ipy = ((float)y + jy) / (float)h * 2.0f - 1.0f;
    dir = normalize(ipx * cam.U + ipy * cam.V + cam.W);
    return rng;
}

ORF_HD f3 path_radiance(const SceneGPU& sc, f3 o, f3 d, uint32_t& rng, unsigned* stk, int stride, PathStats& st)
{
    f3 L = mk3(0.0f), thr = mk3(1.0f);
    int depth = 0, emit = 1;
    float prev_bsdf_pdf = -1.0f;  // > 0: the ray left a Lambert vertex that sampled lights (env MIS)
    const float eps = sc.scene_epsilon;
    for (;;) {
        Hit h;
        st.ext++;
        if (!trace<false>(sc, o, d, eps, kRayMax, h, stk, stride)) {
            if (sc.has_env) {
                f3 E = env_lookup(sc.env, d);
                if (prev_bsdf_pdf > 0.0f && sc.env.sampled)
                    E = E * mis_power(prev_bsdf_pdf, env_pdf(sc.env, d) / (float)sc.nlights);
                L += thr * E;
            }
            break;
        }
        f3 p = o + h.t * d;
        const float4* tg = sc.tri + 3 * (size_t)h.leaf;
        float4 q1 = tg[1], q2 = tg[2];
        int4 sh;
        f3 normal = shading_normal(sc, h.prim, h.b, h.g, mk3(q1.x, q1.y, q1.z), mk3(q2.x, q2.y, q2.z), sh);
        f3 ffn = faceforward(normal, -d, normal);
        const MaterialGPU& mat = sc.materials[sh.w];
        if (mat.kind == MAT_EMITTER) {
            if (dot(normal, -d) > 0.0f && emit) L += thr * mat.radiance;
            break;
        }
        if (mat.kind == MAT_LAMBERT) {
            if (depth > sc.max_depth) break;
            f3 rho = mat.diffuse_color;
            unsigned li = (unsigned)((float)sc.nlights * rnd_tea(rng));
            if (sc.nlights > 0) {
                if (li >= (unsigned)sc.nlights) li = sc.nlights - 1;
                float light_pdf = 1.0f / (float)sc.nlights;
                LightGPU lt = sc.lights[li];
                f3 wl = mk3(0.0f), rad = mk3(0.0f);
                float dist = 0.0f, penv;
                sample_light(sc, lt, p, rng, wl, rad, dist, penv);
                rad = rad / light_pdf;
                float cs = dot(ffn, wl);
                if (cs > 0.0f) {
                    Hit hs;
                    st.shd++;
                    bool occ = trace<true>(sc, p, wl, eps, dist - eps, hs, stk, stride);
                    f3 Li = rad * (occ ? 0.0f : 1.0f);
                    f3 c = Li * kInvPi * cs * rho;
                    if (penv >= 0.0f) c = c * mis_power(penv * light_pdf, cs * kInvPi);
                    L += thr * c;
                    emit = 0;
                }
            }
            float prob = (rho.x + rho.y + rho.z) / 3.0f;
            float xi = rnd_tea(rng);
            if (!(xi < prob)) break;
            f3 nd = sample_cosine(ffn, rng);
            prev_bsdf_pdf = sc.env.sampled ? dot(ffn, nd) * kInvPi : -1.0f;
            thr = thr * (rho / prob);
            o = p;
            d = nd;
            depth++;
            continue;
        }
        if (mat.kind == MAT_CONDUCTOR) {
            if (depth >= sc.max_depth) break;
            float z1 = rnd_tea(rng);
            float z2 = rnd_tea(rng);
            f3 wo, w;
            conductor_step(-d, ffn, mat, z1, z2, wo, w);
            if (w.x == 0.0f && w.y == 0.0f && w.z == 0.0f) break;
            thr = thr * w;
            o = p;
            d = wo;
            depth++;
            emit = 1;
            prev_bsdf_pdf = -1.0f;
            continue;
        }
        if (mat.kind == MAT_DIELECTRIC) {
            if (depth >= sc.max_depth) break;
            f3 nd, f;
            if (!dielectric_step(d, h.t, normal, ffn, mat, rng, nd, f)) break;
            thr = thr * f;
            o = p;
            d = nd;
            depth++;
            emit = 1;
            prev_bsdf_pdf = -1.0f;
            continue;
        }
        break;  // materials without a radiance program in this kernel
    }
    return L;
}

}  // namespace orf
