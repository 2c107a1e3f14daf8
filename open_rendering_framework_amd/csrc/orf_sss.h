// orf_sss.h -- diffusion gather of the translucent material (device; also host-callable).
//
// The reference gathers a translucent hit from every probe of the frame
// (CU/subsurface_scattering_shader.cu:120-167: sum over N probes of
// T12 L S(d) / exp_term with Russian roulette on exp(-d sigma_tr,min), times
// global_coeff / N).  With 2^20 Poisson-disk probes that O(N) loop is replaced
// by importance sampling of the exit point (King et al. 2013, "BSSRDF importance
// sampling", one projection axis):
//   * channel c uniform, radius r ~ sigma_tr,c exp(-sigma_tr,c r), angle uniform
//     on the tangent plane of the entry point xo;
//   * a probe ray along -no through that plane point, clipped to the sphere of
//     radius r_max, collects in one traversal the first kSssHits crossings of
//     the same translucent material; one is picked uniformly;
//   * the exit point's incident light is the nearest probe's (within the grid
//     cell size), its pdf on the surface is
//         p = (1/3) sum_c sigma_c exp(-sigma_c r) / (2 pi r) * |n_i . no| / hits,
//   * estimate = global_coeff * mean over samples of T12 Le S(xi, xo) / p.
// This is the reference's integral (its L already carries the 1 / pdf of the
// probe position) with a different Monte-Carlo estimator.
#pragma once

#include "orf_dipole.h"
#include "orf_sss_grad.h"
#include "orf_shade.h"
#include "orf_trace.h"

namespace orf {

constexpr int kSssHits = 4;            // crossings kept per probe ray
constexpr float kSssRadiusSigmas = 12.0f;  // r_max = 12 / sigma_tr,min

ORF_HD unsigned probe_hash(int ix, int iy, int iz, unsigned mask)
{
    return ((unsigned)ix * 73856093u ^ (unsigned)iy * 19349663u ^ (unsigned)iz * 83492791u) & mask;
}

ORF_HD int probe_cell(float v, float inv_cell) { return (int)floorf(v * inv_cell); }

// Nearest probe strictly within distance `cell` of y (ties: lower index), or -1:
// the smallest (d^2, index) over the probes of the 27 cells around y.  The home
// cell is searched first; a neighbour cell is skipped when its box lies farther
// from y than the best probe so far, since none of its probes can then be the
// answer or tie with it.  (Visiting y's nearer neighbours next, by a per-lane
// mirrored order, measured 3% slower.)  The box distance is shortened by a
// margin that covers the rounding of the cell coordinates (v * inv, 2^-18
// (|v * inv| + 1) cell units) and of the probes' d^2 (a 2^-16 relative margin),
// so the skip is conservative and the result is the full search's.
ORF_HD int nearest_probe(const ProbesGPU& pr, f3 y)
{
    const float inv = 1.0f / pr.cell;
    const float ax = y.x * inv, ay = y.y * inv, az = y.z * inv;
    const int cx = probe_cell(y.x, inv), cy = probe_cell(y.y, inv), cz = probe_cell(y.z, inv);
    // gap from y to the lower / upper neighbour along each axis, in cell units
    auto gap_lo = [](float a, int c) { return fmaxf(a - (float)c - 3.814697265625e-6f * (fabsf(a) + 1.0f), 0.0f); };
    auto gap_hi = [](float a, int c) {
        return fmaxf((float)c + 1.0f - a - 3.814697265625e-6f * (fabsf(a) + 1.0f), 0.0f);
    };
    const float lx = gap_lo(ax, cx), hx = gap_hi(ax, cx), ly = gap_lo(ay, cy), hy = gap_hi(ay, cy),
                lz = gap_lo(az, cz), hz = gap_hi(az, cz);
    const float cell2 = pr.cell * pr.cell;
    const float scale = cell2 * 0.9999847412109375f;  // cell^2 (1 - 2^-16)
    float best_d2 = cell2;
    int best = -1;
    for (int i = 0; i < 27; ++i) {
        const int k = i < 14 ? i + 13 : i - 14;  // 13 (the home cell) first
        const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
        const float gx = dx < 0 ? lx : (dx > 0 ? hx : 0.0f);
        const float gy = dy < 0 ? ly : (dy > 0 ? hy : 0.0f);
        const float gz = dz < 0 ? lz : (dz > 0 ? hz : 0.0f);
        if ((gx * gx + gy * gy + gz * gz) * scale > best_d2) continue;
        uint2 c = pr.cells[probe_hash(cx + dx, cy + dy, cz + dz, pr.table_mask)];
        for (unsigned j = c.x; j < c.x + c.y; ++j) {
            float4 q = pr.pos[j];
            f3 e = mk3(q.x, q.y, q.z) - y;
            float d2 = dot(e, e);
            if (d2 < best_d2 || (d2 == best_d2 && best >= 0 && (int)j < best)) {
                best_d2 = d2;
                best = (int)j;
            }
        }
    }
    return best;
}

struct SssHit {
    float t;
    int prim, leaf;
    float b, g;
};

// Keeps out[0..n) sorted by t with distinct t values, at most kSssHits: a
// crossing at an existing t replaces it when its triangle id is lower.
ORF_HD void sss_insert(SssHit* out, int& n, const SssHit& c)
{
    for (int i = 0; i < n; ++i)
        if (out[i].t == c.t) {
            if (c.prim < out[i].prim) out[i] = c;
            return;
        }
    if (n == kSssHits && !(c.t < out[n - 1].t)) return;
    int i = n < kSssHits ? n++ : kSssHits - 1;
    while (i > 0 && out[i - 1].t > c.t) {
        out[i] = out[i - 1];
        --i;
    }
    out[i] = c;
}

// Crossings of the translucent material `sss` along o + t dir, 0 < t < tmax, in
// one BVH traversal: the kSssHits smallest distinct t, ascending (at equal t
// the lower triangle id), whatever the traversal order.  Once the list is full,
// boxes and triangles beyond its last t (plus one ulp, so equal-t ties still
// resolve) are culled.
ORF_HD int probe_ray_hits(const SceneGPU& sc, int sss, f3 o, f3 dir, float tmax, SssHit* out, unsigned* stk,
                          int stride)
{
    RayPre r = ray_pre(o, dir);
    int n = 0;
    float tcull = tmax;
    unsigned node = 0;
    int sp = 0;
    unsigned ovf[kStackDepth - kLdsStack + 1];
    auto push = [&](unsigned v) { stack_push(stk, stride, ovf, sp, v); };
    auto pop = [&]() -> unsigned { return stack_pop(stk, stride, ovf, sp); };
    const float inf = __builtin_huge_valf();
    for (;;) {
        if (!link_is_leaf(node)) {
            // 4-wide node: entered children near to far, the nearest next, the
            // others pushed farthest first (near-first culls sooner once the
            // list is full; the list itself does not depend on the order)
            const float4* nd = sc.bvh4 + 8 * (size_t)node;
            const float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5], ln = nd[6];
            float e[4] = {box_enter(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, 0.0f, tcull),
                          box_enter(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, 0.0f, tcull),
                          box_enter(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, 0.0f, tcull),
                          box_enter(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, 0.0f, tcull)};
            unsigned c[4] = {f2u(ln.x), f2u(ln.y), f2u(ln.z), f2u(ln.w)};
            auto order = [&](int a, int b) {
                if (e[b] < e[a]) {
                    const float te = e[a];
                    e[a] = e[b];
                    e[b] = te;
                    const unsigned tc = c[a];
                    c[a] = c[b];
                    c[b] = tc;
                }
            };
            order(0, 1);
            order(2, 3);
            order(0, 2);
            order(1, 3);
            order(1, 2);
            if (e[0] == inf) {
                if (sp == 0) break;
                node = pop();
                continue;
            }
            if (e[3] != inf && sp < kStackDepth) push(c[3]);
            if (e[2] != inf && sp < kStackDepth) push(c[2]);
            if (e[1] != inf && sp < kStackDepth) push(c[1]);
            node = c[0];
            continue;
        }
        unsigned first = leaf_first(node), cnt = leaf_count(node);
        for (unsigned k = 0; k < cnt; ++k) {
            const float4* tp = sc.tri + 3 * (size_t)(first + k);
            float4 q0 = tp[0], q1 = tp[1], q2 = tp[2];
            SssHit c;
            if (!intersect_tri(o, dir, mk3(q0.x, q0.y, q0.z), mk3(q1.x, q1.y, q1.z), mk3(q2.x, q2.y, q2.z), 0.0f,
                               tcull, c.t, c.b, c.g))
                continue;
            c.prim = (int)(f2u(q0.w) & kTriIdMask);
            const MaterialGPU& m = sc.materials[sc.shade[c.prim].w];
            if (m.sss != sss) continue;  // the probes' own translucent material only
            c.leaf = (int)(first + k);
            sss_insert(out, n, c);
            if (n == kSssHits) tcull = orf_u2f(f2u(out[n - 1].t) + 1u);
        }
        if (sp == 0) break;
        node = pop();
    }
    return n;
}

// Diffusion term entering at xo (normal no facing the incident side).  Its random
// numbers come from the path's 64-bit stream, as the reference's gather loop
// draws its roulette numbers (CU/subsurface_scattering_shader.cu:152-156).  The
// probe rays use the caller's traversal stack (stk, stride) -- free between
// traces -- or a private one when stk is null.  With `grad`, the derivatives of
// the returned term with respect to the scaled sigma_a and sigma_s of each
// channel go to grad->dL_a, dL_s (the exit points and their pdf held fixed).
ORF_HD f3 sss_gather(const SceneGPU& sc, const ScatteringGPU& P, int sss, f3 xo, f3 no, Seed64& rng,
                     unsigned* stk = nullptr, int stride = 1, SssGrad* grad = nullptr)
{
    const ProbesGPU& pr = sc.probes;
    f3 acc = mk3(0.0f);
    if (pr.count <= 0 || pr.samples <= 0) return acc;
    unsigned own[kLdsStack];
    if (!stk) {
        stk = own;
        stride = 1;
    }
    f3 t1, t2;
    make_frame(no, t1, t2);
    float smin = fminf(P.transport.x, fminf(P.transport.y, P.transport.z));
    float r_max = kSssRadiusSigmas / smin;
    for (int k = 0; k < pr.samples; ++k) {
        const float u = rnd_accurate(rng), v = rnd_accurate(rng), w = rnd_accurate(rng), s = rnd_accurate(rng);
        int c = min(2, (int)(3.0f * u));
        float sig = c == 0 ? P.transport.x : (c == 1 ? P.transport.y : P.transport.z);
        float r = -orf_logf(1.0f - v) / sig;
        if (!(r < r_max)) continue;
        float sp, cp;
        orf_sincosf(6.28318530717958647692f * w, &sp, &cp);
        float hh = sqrtf(r_max * r_max - r * r);
        f3 top = xo + (t1 * cp + t2 * sp) * r + no * hh;
        SssHit hits[kSssHits];
        int nh = probe_ray_hits(sc, sss, top, -no, 2.0f * hh, hits, stk, stride);
        if (nh == 0) continue;
        const SssHit& hp = hits[min(nh - 1, (int)(s * (float)nh))];
        f3 xi = top + (-no) * hp.t;
        const float4* tg = sc.tri + 3 * (size_t)hp.leaf;
        float4 q1 = tg[1], q2 = tg[2];
        int4 sh;
        f3 ni = shading_normal(sc, hp.prim, hp.b, hp.g, mk3(q1.x, q1.y, q1.z), mk3(q2.x, q2.y, q2.z), sh);
        float cosn = fabsf(dot(ni, no));
        f3 dx = xi - xo;
        f3 inpl = dx - no * dot(dx, no);
        float rp = sqrtf(dot(inpl, inpl));
        float pr_r = (P.transport.x * orf_expf(-P.transport.x * rp) + P.transport.y * orf_expf(-P.transport.y * rp) +
                      P.transport.z * orf_expf(-P.transport.z * rp)) /
                     (3.0f * 6.28318530717958647692f * rp);
        float pdf = pr_r * cosn / (float)nh;
        if (!(pdf > 0.0f) || !__builtin_isfinite(pdf)) continue;
        int j = nearest_probe(pr, xi);
        if (j < 0) continue;
        float4 Lj = pr.L[j], wj = pr.w12[j];
        f3 E = mk3(Lj.x, Lj.y, Lj.z) * wj.w;
        f3 S;
        if (P.dipole_model == 1) {
            float4 nj = pr.sn[j];
            S = dirpole_bssrdf(xi, mk3(nj.x, nj.y, nj.z), mk3(wj.x, wj.y, wj.z), xo, no, P);
        } else {
            S = dipole_bssrdf(sqrtf(dot(dx, dx)), P);
        }
        acc += E * S / pdf;
        if (grad) {
            const float4 nj = pr.sn[j];
            const float e[3] = {E.x, E.y, E.z};
            const float sa[3] = {P.absorption.x, P.absorption.y, P.absorption.z};
            const float ss[3] = {P.scattering.x, P.scattering.y, P.scattering.z};
            const float g[3] = {P.meancosine.x, P.meancosine.y, P.meancosine.z};
            float da[3], ds[3];
            for (int c = 0; c < 3; ++c) {
                const SssCoefD cd = sss_coef_channel(sa[c], ss[c], g[c], P.relative_ior);
                const Dual2 Sd = P.dipole_model == 1
                                     ? dirpole_channel(xi, mk3(nj.x, nj.y, nj.z), mk3(wj.x, wj.y, wj.z), xo, no, cd)
                                     : dipole_channel(sqrtf(dot(dx, dx)), cd);
                da[c] = e[c] * Sd.a / pdf;
                ds[c] = e[c] * Sd.s / pdf;
            }
            grad->dL_a += mk3(da[0], da[1], da[2]);
            grad->dL_s += mk3(ds[0], ds[1], ds[2]);
        }
    }
    if (grad) {
        grad->dL_a = grad->dL_a * (1.0f / (float)pr.samples) * P.global_coeff;
        grad->dL_s = grad->dL_s * (1.0f / (float)pr.samples) * P.global_coeff;
    }
    return acc * (1.0f / (float)pr.samples) * P.global_coeff;
}

}  // namespace orf
