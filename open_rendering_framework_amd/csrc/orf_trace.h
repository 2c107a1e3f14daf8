// orf_trace.h -- ray/triangle intersection and 4-wide BVH traversal (device).
//
// Replaces OptiX traversal + intersect_triangle_earlyexit as used by
// CU/triangle_mesh.cu:46-86 (closest hit for radiance rays, any hit for shadow
// rays whose any-hit programs all terminate: CU/diffuse_shader.cu:33-38,
// CU/metallic_shader.cu:38-43, CU/arealight_shader.cu:27-32).
//
// Triangle test: Moller-Trumbore (1997) on (p0, e1 = p1-p0, e2 = p2-p0), plain
// FP32 (no contraction), so (t, beta, gamma) round exactly as in the CPU
// restatement; beta weights p1, gamma weights p2, the geometric normal is
// e1 x e2.  A hit is accepted when tmin < t < t_best (strict).
// Box test: slabs with precomputed 1/d and o/d (explicit fma), far distance
// inflated by 1 + 2^-21 so rounding never culls a box containing a hit.
// Disk and sphere emitters are tested one by one after the BVH (a scene has a
// handful; none in most scenes).
// Traversal stack: the top kLdsStack entries of a lane live in a column of the
// workgroup's LDS array (stride = block size), deeper ones (up to kStackDepth)
// in a private array (the context refuses a tree whose 4-wide traversal could
// hold more than kStackDepth entries).
// Top nodes: kernels that copy the first `ntop` 4-wide nodes (the breadth-first
// top levels, kTopFloat4 float4 each) into LDS pass that copy; a lane at such a
// node reads it from LDS, so the most visited nodes cost no vector-memory
// instructions (the texture address/data path is what bounds the traversal).
// Only where the node's bytes come from changes, never the visit order.
#pragma once

#include "orf_scene.h"

namespace orf {

// Pointer to float4 data in LDS.  A generic pointer to the LDS copy of the top
// nodes compiles to flat loads, which still pass through the texture address
// unit; the LDS address space gives ds_read_b128.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) const float4 lds_cf4;
#else
typedef const float4 lds_cf4;
#endif

struct Hit {
    float t;
    float b, g;   // barycentric weights of p1 and p2
    int prim;     // original triangle id, -1 = miss
    int leaf;     // leaf-order triangle index (geometry record)
};

ORF_HD bool intersect_tri(f3 o, f3 d, f3 p0, f3 e1, f3 e2, float tmin, float tmax, float& t, float& b, float& g)
{
    f3 pv = cross(d, e2);
    float det = dot(e1, pv);
    if (det == 0.0f) return false;
    float inv = 1.0f / det;
    f3 tv = o - p0;
    float u = dot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    f3 qv = cross(tv, e1);
    float v = dot(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    float tt = dot(e2, qv) * inv;
    if (!(tt > tmin && tt < tmax)) return false;
    t = tt;
    b = u;
    g = v;
    return true;
}

// Disk emitter (CU/disk_mesh.cu:41-63): the plane is hit when |n.d| >
// scene_epsilon, at t = -(o - c).n / (n.d), inside when |p - c|^2 <= r^2.
ORF_HD bool intersect_disk(f3 o, f3 d, f3 c, f3 n, float r, float eps, float tmin, float tmax, float& t)
{
    f3 O = o - c;
    float den = dot(n, d);
    if (!(fabsf(den) > eps)) return false;
    float tt = -dot(O, n) / den;
    f3 v = (o + tt * d) - c;
    if (!(dot(v, v) <= r * r)) return false;
    if (!(tt > tmin && tt < tmax)) return false;
    t = tt;
    return true;
}

// Sphere emitter (CU/sphere_mesh.cu:42-91): near root -b - sqrt(b^2 - c) of the
// unit-direction quadratic, refined once from the point it gives when it lies
// beyond 10 radii; the far root is tried when the near one is out of range.
// As in the reference, a refinement whose discriminant is not positive keeps
// the refined b with the first sqrt(disc).
ORF_HD bool intersect_sphere(f3 o, f3 d, f3 c, float r, float tmin, float tmax, float& t)
{
    f3 O = o - c;
    float b = dot(O, d);
    float cc = dot(O, O) - r * r;
    float disc = b * b - cc;
    if (!(disc > 0.0f)) return false;
    float sdisc = sqrtf(disc);
    float root1 = -b - sdisc;
    const bool refine = fabsf(root1) > 10.0f * r;
    float root11 = 0.0f;
    if (refine) {
        f3 O1 = O + root1 * d;
        b = dot(O1, d);
        cc = dot(O1, O1) - r * r;
        disc = b * b - cc;
        if (disc > 0.0f) {
            sdisc = sqrtf(disc);
            root11 = -b - sdisc;
        }
    }
    const float near = root1 + root11;
    if (near > tmin && near < tmax) {
        t = near;
        return true;
    }
    const float root2 = (-b + sdisc) + (refine ? root1 : 0.0f);
    if (root2 > tmin && root2 < tmax) {
        t = root2;
        return true;
    }
    return false;
}

// Analytic emitter record (centre, radius), (disk normal, kind).
ORF_HD bool intersect_analytic(float4 a, float4 b, f3 o, f3 d, float eps, float tmin, float tmax, float& t)
{
    const f3 c = mk3(a.x, a.y, a.z);
    return b.w == 0.0f ? intersect_disk(o, d, c, mk3(b.x, b.y, b.z), a.w, eps, tmin, tmax, t)
                       : intersect_sphere(o, d, c, a.w, tmin, tmax, t);
}

// Unit normal of an analytic emitter at o + t d: the disk normal, or the
// sphere's (o - c + t d) / r (CU/sphere_mesh.cu:73,81), normalised as the
// emitter's closest hit does.
ORF_HD f3 analytic_normal(float4 a, float4 b, f3 o, f3 d, float t)
{
    if (b.w == 0.0f) return normalize(mk3(b.x, b.y, b.z));
    return normalize(((o - mk3(a.x, a.y, a.z)) + t * d) * (1.0f / a.w));
}

struct RayPre {
    f3 o, d, inv, oinv;
};

ORF_HD RayPre ray_pre(f3 o, f3 d)
{
    RayPre r;
    r.o = o;
    r.d = d;
    r.inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.oinv = mk3(o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z);
    return r;
}

// Entry distance of the slab test, or +inf when the box is missed.
ORF_HD float box_enter(const RayPre& r, float lx, float ly, float lz, float hx, float hy,
                                           float hz, float tmin, float tmax)
{
    float t0x = fmaf(lx, r.inv.x, -r.oinv.x), t1x = fmaf(hx, r.inv.x, -r.oinv.x);
    float t0y = fmaf(ly, r.inv.y, -r.oinv.y), t1y = fmaf(hy, r.inv.y, -r.oinv.y);
    float t0z = fmaf(lz, r.inv.z, -r.oinv.z), t1z = fmaf(hz, r.inv.z, -r.oinv.z);
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
    tf *= 1.00000048f;
    return tn <= tf ? tn : __builtin_huge_valf();
}

// Traversal stack: entry i lives in LDS slot i mod kLdsStack while it is among
// the top kLdsStack entries; a push that reuses a slot first moves the deeper
// entry it holds to the private array `ovf`, and the pop that frees the slot
// moves it back.  Every pop and push is an LDS access; the private array is
// touched only past kLdsStack entries.
ORF_HD void stack_push(unsigned* stk, int stride, unsigned* ovf, int& sp, unsigned v)
{
    const int slot = stride * (sp & (kLdsStack - 1));
    if (sp >= kLdsStack) ovf[sp - kLdsStack] = stk[slot];
    stk[slot] = v;
    ++sp;
}

ORF_HD unsigned stack_pop(unsigned* stk, int stride, const unsigned* ovf, int& sp)
{
    --sp;
    const int slot = stride * (sp & (kLdsStack - 1));
    const unsigned v = stk[slot];
    if (sp >= kLdsStack) stk[slot] = ovf[sp - kLdsStack];
    return v;
}

// Closest hit (any-hit off) or first hit (any-hit on) in (tmin, tmax).  kMode 0:
// closest, 1: any, 2: chosen per lane by `any` (one traversal loop serves
// extension and shadow rays of different lanes at once).  kEmit: the disk and
// sphere emitters are tested; kPass: shadow rays pass through triangles marked
// kTriNoShadow (kernels for scenes that cannot hold either leave it out).
template <int kMode, bool kEmit = true, bool kPass = true, bool kTop = false>
ORF_HD bool trace_impl(const SceneGPU& sc, f3 o, f3 d, float tmin, float tmax, bool any, Hit& h, unsigned* stk,
                       int stride, lds_cf4* top = nullptr, unsigned ntop = 0u)
{
    const bool ANY = kMode == 1 || (kMode == 2 && any);
    RayPre r = ray_pre(o, d);
    float tbest = tmax;
    int prim = -1, leaf = -1;
    float hb = 0.0f, hg = 0.0f;
    unsigned node = 0;
    int sp = 0;
    unsigned ovf[kStackDepth - kLdsStack + 1];
    auto push = [&](unsigned v) { stack_push(stk, stride, ovf, sp, v); };
    auto pop = [&]() -> unsigned { return stack_pop(stk, stride, ovf, sp); };
    bool have = true;  // `node` holds a live internal node or leaf link
    while (have) {
        if (!link_is_leaf(node)) {
            // 4-wide node: the entered children in near-to-far order; the
            // nearest is taken next, the others pushed farthest first
            float4 lx, hx, ly, hy, lz, hz, ln;
            if (kTop && node < ntop) {
                lds_cf4* nd = top + kTopFloat4 * node;
                lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5], ln = nd[6];
                // keeps the two arms apart: merged, they become flat loads through phi'd pointers
                __asm__ volatile("" ::: "memory");
            } else {
                const float4* nd = sc.bvh4 + 8 * (size_t)node;
                lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5], ln = nd[6];
            }
            float e[4] = {box_enter(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, tmin, tbest),
                          box_enter(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, tmin, tbest),
                          box_enter(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, tmin, tbest),
                          box_enter(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, tmin, tbest)};
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
            const float inf = __builtin_huge_valf();
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
            float t, bb, gg;
            if (intersect_tri(o, d, mk3(q0.x, q0.y, q0.z), mk3(q1.x, q1.y, q1.z), mk3(q2.x, q2.y, q2.z), tmin,
                              tbest, t, bb, gg)) {
                const unsigned id = f2u(q0.w);
                if (ANY && kPass && (id & kTriNoShadow)) continue;  // shadow rays pass through it
                tbest = t;
                hb = bb;
                hg = gg;
                prim = (int)(id & kTriIdMask);
                leaf = (int)(first + k);
                if (ANY) {
                    h.t = t;
                    h.b = bb;
                    h.g = gg;
                    h.prim = prim;
                    h.leaf = leaf;
                    return true;
                }
            }
        }
        if (sp == 0) break;
        node = pop();
    }
    for (int i = 0; kEmit && i < sc.nanalytic; ++i) {
        float t;
        if (intersect_analytic(sc.analytic[2 * i], sc.analytic[2 * i + 1], o, d, sc.scene_epsilon, tmin, tbest, t)) {
            tbest = t;
            prim = sc.analytic_prim + i;
            leaf = -1 - i;
            if (ANY) break;
        }
    }
    h.t = tbest;
    h.b = hb;
    h.g = hg;
    h.prim = prim;
    h.leaf = leaf;
    return prim >= 0;
}

template <bool ANY, bool kEmit = true, bool kPass = true>
ORF_HD bool trace(const SceneGPU& sc, f3 o, f3 d, float tmin, float tmax, Hit& h, unsigned* stk, int stride)
{
    return trace_impl<ANY ? 1 : 0, kEmit, kPass>(sc, o, d, tmin, tmax, ANY, h, stk, stride);
}

// As trace, reading the first ntop nodes from the LDS copy `top`.
template <bool ANY, bool kEmit = true, bool kPass = true>
ORF_HD bool trace_top(const SceneGPU& sc, f3 o, f3 d, float tmin, float tmax, Hit& h, unsigned* stk, int stride,
                      lds_cf4* top, unsigned ntop)
{
    return trace_impl<ANY ? 1 : 0, kEmit, kPass, true>(sc, o, d, tmin, tmax, ANY, h, stk, stride, top, ntop);
}

// Cooperative copy of the first min(nodes4, cap) nodes into the LDS array `top`
// by the whole workgroup (every thread calls it; ends with a barrier).
__device__ __forceinline__ unsigned load_top_nodes(const SceneGPU& sc, float4* top, unsigned cap)
{
    const unsigned n = min((unsigned)sc.nodes4, min(cap, (unsigned)kBvhTopOrdered));
    for (unsigned i = threadIdx.x; i < n * kTopFloat4; i += blockDim.x) {
        const unsigned node = i / kTopFloat4, row = i - node * kTopFloat4;
        top[i] = sc.bvh4[8 * (size_t)node + row];
    }
    __syncthreads();
    return n;
}

ORF_HD bool trace_dyn(const SceneGPU& sc, f3 o, f3 d, float tmin, float tmax, bool any, Hit& h, unsigned* stk,
                      int stride)
{
    return trace_impl<2>(sc, o, d, tmin, tmax, any, h, stk, stride);
}

}  // namespace orf
