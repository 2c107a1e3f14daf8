/*
 * oracle_render.c -- CPU restatement of the per-pixel render path (TEST INFRASTRUCTURE).
 *
 * Recursive like the reference programs (each closest hit traces its child and
 * adds child * weight), so the summation order is the reference's; the product
 * kernels iterate with a throughput instead.  The acceleration structure is the
 * oracle's own object-median BVH with a double-precision, slightly inflated box
 * test; the triangle test is Moller-Trumbore in R, the same arithmetic as the
 * product's, so hits agree to the bit in the FP32 oracle.
 *
 * New parts with no reference counterpart (marked where they occur): the
 * environment as a sampled light with power-heuristic MIS (the reference's env
 * importance sampling, CU/env_cameras.cu + CU/envmap_shader.cu, is never
 * compiled), clamping of light picks where rnd_tea rounds to 1.0f, and the
 * film as sums of finite samples.
 */
#include "oracle_render.h"

#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

#include "oracle_fmath.h"

#define PI_R ((R)3.14159265358979323846)
#define INV_PI_R ((R)0.318309886183790671538)
#define RAY_MAX ((R)1e27f)
#define MS_WALK_CAP 10 /* scattering events per microsurface walk (new: the reference's loop is unbounded) */
#define MS_DIFFUSE_CAP 100 /* scattering events per rough diffuse walk (the reference's own bound on its sampler) */

typedef struct {
    double lo[3], hi[3];
    int32_t left, right;  /* left < 0: leaf */
    int32_t first, count;
} bnode;

/* ScatteringMaterialProperties (ORF/structs.h:116-144) in R. */
typedef struct {
    R relative_ior;
    int dipole_model;
    v3 absorption, scattering, meancosine, extinction, reducedExtinction, deltaEddExtinction, D, transport;
    v3 C_phi, C_phi_inv, C_E, A, albedo, reducedAlbedo, de, three_D, two_A_de, global_coeff, one_over_three_ext;
    R mean_transport;
} osss;

typedef struct {
    /* geometry (world space) */
    int64_t ntris;
    v3* p0;
    v3* e1;
    v3* e2;
    int32_t* nidx;  /* 3 per tri or NULL */
    v3* normals;
    int32_t* tri_mat;
    /* bvh */
    bnode* nodes;
    int32_t nnodes;
    int32_t* order;
    /* materials and lights */
    int32_t nmats;
    orc_material* mats;
    cplx3* eta;
    cplx3* ieta; /* 1 / eta = the complex ior (conductor walk events met from below) */
    int32_t nlights;
    orc_light* lights;
    orc_trilight* trilights;
    /* environment */
    float* env;
    int32_t env_w, env_h, env_is, has_env;
    R* mpdf;
    R* mcdf;
    R* cpdf;
    R* ccdf;
    /* camera and integrator */
    v3 eye, U, V, W;
    int32_t max_depth;
    R eps;
    /* differentiable form (new, no reference counterpart): textured Lambert albedo
     * and GGX-conductor roughness; *_eval shade, *_frozen drive sampling and RR */
    int32_t diff_on, tex_w, tex_h, alb_mat, rough_mat, alb_uv, rough_uv;
    R alb_uvp[4], rough_uvp[4];
    R* alb_eval;
    R* alb_frozen;
    R* rough_eval;
    R* rough_frozen;
    /* translucent materials: coefficients per material, probes */
    osss* sss;
    /* differentiable form (new): material sss_diff_mat shades with sss[] (the
     * evaluated coefficients) and samples -- gather exit points, extinction
     * roulette -- with sss_frozen */
    int32_t sss_diff_mat;
    osss sss_frozen;
    int32_t nprobes, probe_samples, probe_mat;
    R probe_cell;
    v3* probe_pos;
    v3* probe_nrm;
    v3* probe_L;
    v3* probe_w12;
    R* probe_T12;
    v3* probe_sn; /* normal the probe's light entered through (microfacet normal of a rough boundary) */
    /* disk and sphere emitters (one per such light, in light order); a hit on
     * emitter i reports prim ntris + i */
    int32_t nan;
    v3* an_c;
    v3* an_n; /* disk normal */
    R* an_r;
    int32_t* an_sphere;
} oscene;

typedef struct {
    v3 eye, U, V, W;
} ocam;

/* ---------------------------------------------------------------- BVH */

typedef struct {
    double c[3];
    int32_t id;
} cref;


static void cref_swap(cref* a, cref* b)
{
    cref t = *a;
    *a = *b;
    *b = t;
}

/* quickselect: place the k-th smallest (by c[axis]) at position k */
static void select_k(cref* a, int64_t n, int64_t k, int axis)
{
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        double pivot = a[(lo + hi) / 2].c[axis];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (a[i].c[axis] < pivot) i++;
            while (a[j].c[axis] > pivot) j--;
            if (i <= j) {
                cref_swap(&a[i], &a[j]);
                i++;
                j--;
            }
        }
        if (k <= j) hi = j;
        else if (k >= i) lo = i;
        else break;
    }
}

static void tri_box(const oscene* s, int32_t t, double lo[3], double hi[3])
{
    v3 p[3] = {s->p0[t], vadd(s->p0[t], s->e1[t]), vadd(s->p0[t], s->e2[t])};
    for (int a = 0; a < 3; a++) {
        lo[a] = 1e300;
        hi[a] = -1e300;
    }
    for (int k = 0; k < 3; k++) {
        double q[3] = {p[k].x, p[k].y, p[k].z};
        for (int a = 0; a < 3; a++) {
            if (q[a] < lo[a]) lo[a] = q[a];
            if (q[a] > hi[a]) hi[a] = q[a];
        }
    }
}

static int32_t build_rec(oscene* s, cref* refs, int64_t first, int64_t n, int32_t* cap)
{
    if (s->nnodes >= *cap) {
        *cap *= 2;
        s->nodes = (bnode*)realloc(s->nodes, sizeof(bnode) * (size_t)*cap);
    }
    int32_t me = s->nnodes++;
    bnode nd;
    for (int a = 0; a < 3; a++) {
        nd.lo[a] = 1e300;
        nd.hi[a] = -1e300;
    }
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = first; i < first + n; i++) {
        double lo[3], hi[3];
        tri_box(s, refs[i].id, lo, hi);
        for (int a = 0; a < 3; a++) {
            if (lo[a] < nd.lo[a]) nd.lo[a] = lo[a];
            if (hi[a] > nd.hi[a]) nd.hi[a] = hi[a];
            if (refs[i].c[a] < clo[a]) clo[a] = refs[i].c[a];
            if (refs[i].c[a] > chi[a]) chi[a] = refs[i].c[a];
        }
    }
    /* inflate so a float hit on a box face is never culled */
    for (int a = 0; a < 3; a++) {
        double pad = 1e-6 * (nd.hi[a] - nd.lo[a]) + 1e-6 * (fabs(nd.lo[a]) + fabs(nd.hi[a])) + 1e-30;
        nd.lo[a] -= pad;
        nd.hi[a] += pad;
    }
    if (n <= 4) {
        nd.left = nd.right = -1;
        nd.first = (int32_t)first;
        nd.count = (int32_t)n;
        s->nodes[me] = nd;
        return me;
    }
    int axis = 0;
    for (int a = 1; a < 3; a++)
        if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
    int64_t half = n / 2;
    select_k(refs + first, n, half, axis);
    nd.first = nd.count = 0;
    s->nodes[me] = nd;
    int32_t l = build_rec(s, refs, first, half, cap);
    int32_t r = build_rec(s, refs, first + half, n - half, cap);
    s->nodes[me].left = l;
    s->nodes[me].right = r;
    return me;
}

static void build_bvh(oscene* s)
{
    int64_t n = 0;
    cref* refs = (cref*)malloc(sizeof(cref) * (size_t)(s->ntris + 1));
    for (int64_t t = 0; t < s->ntris; t++) {
        /* mesh_bounds: triangles of zero or non-finite area get no box (CU/triangle_mesh.cu:88-105) */
        R area = sqrt(vdot(vcross(s->e1[t], s->e2[t]), vcross(s->e1[t], s->e2[t])));
        float af = (float)area;
        if (!(af > 0.0f) || isinf(af)) continue;
        double lo[3], hi[3];
        tri_box(s, (int32_t)t, lo, hi);
        for (int a = 0; a < 3; a++) refs[n].c[a] = 0.5 * (lo[a] + hi[a]);
        refs[n].id = (int32_t)t;
        n++;
    }
    int32_t cap = 1024;
    s->nodes = (bnode*)malloc(sizeof(bnode) * (size_t)cap);
    s->nnodes = 0;
    if (n > 0) build_rec(s, refs, 0, n, &cap);
    s->order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < n; i++) s->order[i] = refs[i].id;
    free(refs);
}

static int box_hit(const bnode* b, const double o[3], const double inv[3], double tmin, double tmax, double* tn)
{
    double t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; a++) {
        double ta = (b->lo[a] - o[a]) * inv[a], tb = (b->hi[a] - o[a]) * inv[a];
        if (ta != ta) ta = -1e300; /* 0 * inf on a box face: no constraint */
        if (tb != tb) tb = 1e300;
        if (ta > tb) {
            double x = ta;
            ta = tb;
            tb = x;
        }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
    }
    *tn = t0;
    return t0 <= t1;
}

/* Moller-Trumbore (1997), same operation order as the product's kernels. */
static int isect(v3 o, v3 d, v3 p0, v3 e1, v3 e2, R tmin, R tmax, R* t, R* b, R* g)
{
    v3 pv = vcross(d, e2);
    R det = vdot(e1, pv);
    if (det == 0) return 0;
    R inv = (R)1 / det;
    v3 tv = vsub(o, p0);
    R u = vdot(tv, pv) * inv;
    if (u < 0 || u > 1) return 0;
    v3 qv = vcross(tv, e1);
    R v = vdot(d, qv) * inv;
    if (v < 0 || u + v > 1) return 0;
    R tt = vdot(e2, qv) * inv;
    if (!(tt > tmin && tt < tmax)) return 0;
    *t = tt;
    *b = u;
    *g = v;
    return 1;
}

typedef struct {
    R t, b, g;
    int32_t prim;
    v3 n; /* unit normal of a disk or sphere emitter hit */
} ohit;

/* Disk emitter (CU/disk_mesh.cu:41-63): plane hit when |n.d| > scene_epsilon, at
 * t = -(o - c).n / (n.d), kept when |p - c|^2 <= r^2. */
static int isect_disk(v3 o, v3 d, v3 c, v3 n, R r, R eps, R tmin, R tmax, R* t)
{
    v3 O = vsub(o, c);
    R den = vdot(n, d);
    if (!(fabs(den) > eps)) return 0;
    R tt = -vdot(O, n) / den;
    v3 v = vsub(vadd(o, vscale(d, tt)), c);
    if (!(vdot(v, v) <= r * r)) return 0;
    if (!(tt > tmin && tt < tmax)) return 0;
    *t = tt;
    return 1;
}

/* Sphere emitter (CU/sphere_mesh.cu:42-91): near root -b - sqrt(b^2 - c), refined
 * once from its point when beyond 10 radii (a refinement with disc <= 0 keeps
 * the refined b and the first sqrt); the far root when the near one is out of
 * (tmin, tmax). */
static int isect_sphere(v3 o, v3 d, v3 c, R r, R tmin, R tmax, R* t)
{
    v3 O = vsub(o, c);
    R b = vdot(O, d);
    R cc = vdot(O, O) - r * r;
    R disc = b * b - cc;
    if (!(disc > 0)) return 0;
    R sdisc = sqrt(disc);
    R root1 = -b - sdisc;
    int refine = fabs(root1) > (R)10.0f * r;
    R root11 = 0;
    if (refine) {
        v3 O1 = vadd(O, vscale(d, root1));
        b = vdot(O1, d);
        cc = vdot(O1, O1) - r * r;
        disc = b * b - cc;
        if (disc > 0) {
            sdisc = sqrt(disc);
            root11 = -b - sdisc;
        }
    }
    R nr = root1 + root11;
    if (nr > tmin && nr < tmax) {
        *t = nr;
        return 1;
    }
    R root2 = (-b + sdisc) + (refine ? root1 : (R)0);
    if (root2 > tmin && root2 < tmax) {
        *t = root2;
        return 1;
    }
    return 0;
}

/* Emitters after the triangles: closest (or any) hit among them in (tmin, h->t). */
static int trace_emitters(const oscene* s, v3 o, v3 d, R tmin, int any, ohit* h)
{
    for (int32_t i = 0; i < s->nan; i++) {
        R t;
        int hit = s->an_sphere[i] ? isect_sphere(o, d, s->an_c[i], s->an_r[i], tmin, h->t, &t)
                                  : isect_disk(o, d, s->an_c[i], s->an_n[i], s->an_r[i], s->eps, tmin, h->t, &t);
        if (!hit) continue;
        h->t = t;
        h->b = h->g = 0;
        h->prim = (int32_t)s->ntris + i;
        /* normalize(shading_normal): the disk normal, or (o - c + t d) / r */
        h->n = s->an_sphere[i] ? orc_normalize(vscale(vadd(vsub(o, s->an_c[i]), vscale(d, t)), (R)1.0f / s->an_r[i]))
                               : orc_normalize(s->an_n[i]);
        if (any) return 1;
    }
    return h->prim >= 0;
}

static int trace(const oscene* s, v3 o, v3 d, R tmin, R tmax, int any, ohit* h)
{
    h->prim = -1;
    h->t = tmax;
    if (s->nnodes == 0) return trace_emitters(s, o, d, tmin, any, h);
    double od[3] = {o.x, o.y, o.z};
    double inv[3] = {1.0 / (double)d.x, 1.0 / (double)d.y, 1.0 / (double)d.z};
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const bnode* b = &s->nodes[stack[--sp]];
        double tn;
        if (!box_hit(b, od, inv, (double)tmin, (double)h->t * (1.0 + 1e-6), &tn)) continue;
        if (b->left < 0) {
            for (int32_t k = 0; k < b->count; k++) {
                int32_t id = s->order[b->first + k];
                R t, bb, gg;
                if (isect(o, d, s->p0[id], s->e1[id], s->e2[id], tmin, h->t, &t, &bb, &gg)) {
                    /* normal_shader's any-hit terminates without clearing the attenuation
                     * (CU/normal_shader.cu:18-22): shadow rays pass through it */
                    if (any && s->mats[s->tri_mat[id]].kind == ORC_NORMAL) continue;
                    h->t = t;
                    h->b = bb;
                    h->g = gg;
                    h->prim = id;
                    if (any) return 1;
                }
            }
            continue;
        }
        stack[sp++] = b->right;
        stack[sp++] = b->left;
    }
    return trace_emitters(s, o, d, tmin, any, h);
}

static int trace_brute(const oscene* s, v3 o, v3 d, R tmin, R tmax, ohit* h)
{
    h->prim = -1;
    h->t = tmax;
    for (int64_t id = 0; id < s->ntris; id++) {
        R t, bb, gg;
        R area = sqrt(vdot(vcross(s->e1[id], s->e2[id]), vcross(s->e1[id], s->e2[id])));
        if (!((float)area > 0.0f) || isinf((float)area)) continue;
        if (isect(o, d, s->p0[id], s->e1[id], s->e2[id], tmin, h->t, &t, &bb, &gg)) {
            h->t = t;
            h->b = bb;
            h->g = gg;
            h->prim = (int32_t)id;
        }
    }
    return trace_emitters(s, o, d, tmin, 0, h);
}

/* ---------------------------------------------------------------- environment */

/* tex2D, normalized coordinates, repeat wrap, linear filter (ORF/HDRLoader.cpp:231-286). */
static v3 tex_bilinear(const oscene* s, R u, R v)
{
    int w = s->env_w, h = s->env_h;
    R x = u * (R)w - (R)0.5f, y = v * (R)h - (R)0.5f;
    R fx0 = floor(x), fy0 = floor(y);
    R fx = x - fx0, fy = y - fy0;
    int ix = (int)fx0 % w, iy = (int)fy0 % h;
    if (ix < 0) ix += w;
    if (iy < 0) iy += h;
    int ix1 = ix + 1 == w ? 0 : ix + 1, iy1 = iy + 1 == h ? 0 : iy + 1;
    const float* t00 = s->env + 4 * ((size_t)iy * w + ix);
    const float* t10 = s->env + 4 * ((size_t)iy * w + ix1);
    const float* t01 = s->env + 4 * ((size_t)iy1 * w + ix);
    const float* t11 = s->env + 4 * ((size_t)iy1 * w + ix1);
    R gx = (R)1 - fx, gy = (R)1 - fy;
    R r[3];
    for (int k = 0; k < 3; k++) {
        R a = (R)t00[k] * gx + (R)t10[k] * fx;
        R b = (R)t01[k] * gx + (R)t11[k] * fx;
        r[k] = a * gy + b * fy;
    }
    return V3(r[0], r[1], r[2]);
}

/* env_lookup (ORF/Envmap.h:57-64). */
static v3 env_lookup(const oscene* s, v3 dir)
{
    R theta = F_ACOS(dir.y);
    R phi = F_ATAN2(dir.x, dir.z);
    R u = (phi + PI_R) * (R)0.5f * INV_PI_R;
    R v = (R)(1.0 - (double)(theta * INV_PI_R));
    return tex_bilinear(s, u, v);
}

/* cdf_bsearch_marginal / cdf_bsearch_conditional (ORF/Envmap.h:23-55) pick the
 * cell i with cdf[i-1] <= xi <= cdf[i].  Where xi equals the value of a flat
 * run of the cdf (always so for xi = 0 when the first cells have zero density)
 * that bisection can land on a zero-width cell and the sampler then divides 0/0.
 * Deviation: the first i with cdf[i] > xi (n - 1 if none), which never has zero
 * width; for every other xi it is the same cell. */
static int env_bsearch(const R* cdf, int n, R xi)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] > xi) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

int orc_env_bsearch(const double* cdf, int n, double xi)
{
    R* c = (R*)malloc(sizeof(R) * (size_t)n);
    for (int i = 0; i < n; i++) c[i] = (R)cdf[i];
    int r = env_bsearch(c, n, (R)xi);
    free(c);
    return r;
}

/* env_luminance_camera / env_marginal_camera / env_pdf_camera (CU/env_cameras.cu:18-75)
 * evaluated in double; every row's last conditional cdf entry and the last
 * marginal entry are set to 1, zero rows get a uniform conditional. */
static void build_env_tables(oscene* s)
{
    int w = s->env_w, h = s->env_h;
    double* rowmean = (double*)malloc(sizeof(double) * (size_t)h);
    s->mpdf = (R*)malloc(sizeof(R) * (size_t)h);
    s->mcdf = (R*)malloc(sizeof(R) * (size_t)h);
    s->cpdf = (R*)malloc(sizeof(R) * (size_t)w * h);
    s->ccdf = (R*)malloc(sizeof(R) * (size_t)w * h);
    double* lum = (double*)malloc(sizeof(double) * (size_t)w);
    for (int j = 0; j < h; j++) {
        double theta = (1.0 - ((double)j + 0.5) / (double)h) * 3.14159265358979323846;
        double st = sin(theta), sum = 0.0;
        for (int i = 0; i < w; i++) {
            const float* t = s->env + 4 * ((size_t)j * w + i);
            lum[i] = ((double)t[0] * 0.2989 + (double)t[1] * 0.5866 + (double)t[2] * 0.1145) * st;
            sum += lum[i];
        }
        double mean = sum / (double)w, run = 0.0;
        rowmean[j] = mean;
        for (int i = 0; i < w; i++) {
            run += lum[i];
            size_t k = (size_t)j * w + i;
            if (mean > 0.0) {
                s->cpdf[k] = (R)(lum[i] / mean);
                s->ccdf[k] = i == w - 1 ? (R)1 : (R)(run / (double)w / mean);
            } else {
                s->cpdf[k] = 1;
                s->ccdf[k] = i == w - 1 ? (R)1 : (R)((double)(i + 1) / (double)w);
            }
        }
    }
    double msum = 0.0;
    for (int j = 0; j < h; j++) msum += rowmean[j];
    double mmean = msum / (double)h, run = 0.0;
    for (int j = 0; j < h; j++) {
        run += rowmean[j];
        s->mpdf[j] = mmean > 0.0 ? (R)(rowmean[j] / mmean) : (R)1;
        s->mcdf[j] = j == h - 1 ? (R)1 : (mmean > 0.0 ? (R)(run / (double)h / mmean) : (R)((j + 1.0) / h));
    }
    free(lum);
    free(rowmean);
}

/* Solid-angle density of env_sample in direction dir (new: MIS). */
static R env_pdf(const oscene* s, v3 dir)
{
    R theta = F_ACOS(dir.y);
    R phi = F_ATAN2(dir.x, dir.z);
    R u = (phi + PI_R) * (R)0.5f * INV_PI_R;
    R v = (R)1 - theta * INV_PI_R;
    int iu = (int)(u * (R)s->env_w), iv = (int)(v * (R)s->env_h);
    if (iu < 0) iu = 0;
    if (iu > s->env_w - 1) iu = s->env_w - 1;
    if (iv < 0) iv = 0;
    if (iv > s->env_h - 1) iv = s->env_h - 1;
    R p = s->mpdf[iv] * s->cpdf[(size_t)iv * s->env_w + iu];
    R st = F_SIN(theta);
    return st > 0 ? p / ((R)2 * PI_R * PI_R * st) : 0;
}

/* sample_environment (ORF/Envmap.h:66-98): returns L = E / pdf and the pdf. */
static void env_sample(const oscene* s, uint32_t* t, v3* dir, v3* L, R* pdf_sa)
{
    const R two_pi2 = (R)2 * PI_R * PI_R;
    R xi1 = orc_rnd_tea(t), xi2 = orc_rnd_tea(t);
    int vi = env_bsearch(s->mcdf, s->env_h, xi1);
    R dv = vi > 0 ? (xi1 - s->mcdf[vi - 1]) / (s->mcdf[vi] - s->mcdf[vi - 1]) : xi1 / s->mcdf[vi];
    R pdf_m = s->mpdf[vi];
    R v = ((R)vi + dv) / (R)s->env_h;
    const R* cc = s->ccdf + (size_t)vi * s->env_w;
    int ui = env_bsearch(cc, s->env_w, xi2);
    R du = ui > 0 ? (xi2 - cc[ui - 1]) / (cc[ui] - cc[ui - 1]) : xi2 / cc[ui];
    R pdf_c = s->cpdf[(size_t)vi * s->env_w + ui];
    R u = ((R)ui + du) / (R)s->env_w;
    R prob = pdf_m * pdf_c;
    R theta = ((R)1 - v) * PI_R;
    R phi = ((R)2 * u - (R)1) * PI_R;
    R st = F_SIN(theta), ct = F_COS(theta), sp = F_SIN(phi), cp = F_COS(phi);
    *dir = V3(st * sp, ct, st * cp);
    v3 E = tex_bilinear(s, u, v);
    *L = vscale(E, st * two_pi2 / prob);
    /* new: a draw that rounds onto a pole (st = 0) has L = 0 and density 0 here
     * (not +inf, which made the MIS weight inf / inf) */
    *pdf_sa = st > 0 ? prob / (two_pi2 * st) : 0;
}

/* ---------------------------------------------------------------- lights */

/* evaluate_direct_illumination (ORF/LightSampler.h:143-166): point (:20-28),
 * directional (:31-37), triangle area light (:40-71), environment (new). */
static void direct_light(const oscene* s, const orc_light* lt, v3 pos, uint32_t* seed, v3* dir, v3* L, R* dist,
                         R* pdf_env)
{
    *pdf_env = -1;
    v3 rad = V3(lt->radiance[0], lt->radiance[1], lt->radiance[2]);
    v3 prm = V3(lt->params[0], lt->params[1], lt->params[2]);
    if (lt->type == ORC_POINT) {
        *dir = vsub(prm, pos);
        R r2 = vdot(*dir, *dir);
        *dist = sqrt(r2);
        *dir = V3(dir->x / *dist, dir->y / *dist, dir->z / *dist);
        *L = V3(rad.x / r2, rad.y / r2, rad.z / r2);
    } else if (lt->type == ORC_DIRECTIONAL) {
        *dir = orc_normalize(vneg(prm));
        *dist = RAY_MAX;
        *L = rad;
    } else if (lt->type == ORC_TRIANGLES) {
        uint32_t ntri = (uint32_t)lt->params[0];
        uint32_t pick = (uint32_t)(orc_rnd_tea(seed) * (R)ntri);
        if (pick >= ntri) pick = ntri - 1; /* new: rnd_tea may round to 1.0f */
        const orc_trilight* tl = &s->trilights[pick + (uint32_t)lt->params[1]];
        v3 uvw = orc_sample_barycentric(seed);
        v3 a = V3(tl->v[0], tl->v[1], tl->v[2]), b = V3(tl->v[3], tl->v[4], tl->v[5]),
           c = V3(tl->v[6], tl->v[7], tl->v[8]);
        v3 lp = vadd(vadd(vscale(a, uvw.x), vscale(b, uvw.y)), vscale(c, uvw.z));
        v3 n;
        if (tl->has_normals) {
            v3 n0 = V3(tl->n[0], tl->n[1], tl->n[2]), n1 = V3(tl->n[3], tl->n[4], tl->n[5]),
               n2 = V3(tl->n[6], tl->n[7], tl->n[8]);
            n = orc_normalize(vadd(vadd(vscale(n0, uvw.x), vscale(n1, uvw.y)), vscale(n2, uvw.z)));
        } else {
            n = orc_normalize(vcross(vsub(b, a), vsub(c, a)));
        }
        *dir = vsub(lp, pos);
        R d2 = vdot(*dir, *dir);
        *dist = sqrt(d2);
        *dir = orc_normalize(*dir);
        R cp = vdot(n, vneg(*dir));
        if (!(cp >= 0)) cp = 0;
        *L = vscale(rad, (R)tl->area * cp / d2 * (R)ntri);
    } else if (lt->type == ORC_DISK) {
        /* evaluate_disk_area_light (ORF/LightSampler.h:73-101) */
        R radius = lt->shape[0];
        R th = (R)lt->shape[1] * PI_R / (R)180.0f, ph = (R)lt->shape[2] * PI_R / (R)180.0f;
        R area = PI_R * radius * radius;
        R pdf = (R)1.0f / area;
        R sth = F_SIN(th), cth = F_COS(th), sph = F_SIN(ph), cph = F_COS(ph);
        v3 n = V3(sth * sph, cth, sth * cph);
        v3 U, Vv;
        orc_create_onb(n, &U, &Vv);
        R u1 = orc_rnd_tea(seed);
        R u2 = orc_rnd_tea(seed);
        R su = sqrt(u1);
        R r = su * radius;
        R t = (R)2.0f * PI_R * u2;
        R dx = r * F_COS(t), dy = r * F_SIN(t), dz = 0;
        v3 dp = vadd(vadd(vscale(U, dx), vscale(Vv, dy)), vscale(n, dz));
        *dir = vsub(vadd(prm, dp), pos);
        R d2 = vdot(*dir, *dir);
        *dist = sqrt(d2);
        *dir = orc_normalize(*dir);
        R cp = vdot(n, vneg(*dir));
        if (!(cp >= 0)) cp = 0;
        R k = cp / d2;
        *L = V3(rad.x * k / pdf, rad.y * k / pdf, rad.z * k / pdf);
    } else if (lt->type == ORC_SPHERE) {
        /* evaluate_spherical_area_light (ORF/LightSampler.h:103-141): the hemisphere
         * facing pos, uniform by area (z = u1), divided by the whole sphere's
         * density 1 / (4 pi R^2) */
        R radius = lt->shape[0];
        v3 vn = orc_normalize(vsub(pos, prm));
        v3 U, Vv;
        orc_create_onb(vn, &U, &Vv);
        R area = (R)4.0f * PI_R * radius * radius;
        R pdf = (R)1.0f / area;
        R u1 = orc_rnd_tea(seed);
        R u2 = orc_rnd_tea(seed);
        R z = u1;
        R sz = sqrt((R)1.0f - z * z);
        R a = (R)2.0f * PI_R * u2;
        R ca = F_COS(a), sa = F_SIN(a);
        v3 sp = V3(ca * sz * radius, sa * sz * radius, z * radius);
        sp = vadd(vadd(vadd(vscale(U, sp.x), vscale(Vv, sp.y)), vscale(vn, sp.z)), prm);
        v3 n = orc_normalize(vsub(sp, prm));
        *dir = vsub(sp, pos);
        R d2 = vdot(*dir, *dir);
        *dist = sqrt(d2);
        *dir = orc_normalize(*dir);
        R cp = vdot(n, vneg(*dir));
        if (!(cp >= 0)) cp = 0;
        R k = cp / d2;
        *L = V3(rad.x * k / pdf, rad.y * k / pdf, rad.z * k / pdf);
    } else if (lt->type == ORC_ENV) {
        env_sample(s, seed, dir, L, pdf_env);
        *dist = RAY_MAX;
    } else {
        *dir = V3(0, 0, 1);
        *L = V3(0, 0, 0);
        *dist = 0;
    }
}

static R mis_power(R a, R b)
{
    R a2 = a * a, b2 = b * b;
    return a2 + b2 > 0 ? a2 / (a2 + b2) : 0;
}

/* ---------------------------------------------------------------- shading */

typedef struct {
    uint64_t ext, shd;
} ostats;

static v3 radiance(const oscene* s, v3 o, v3 d, int depth, int emit_light, uint32_t* seed, R prev_pdf, ostats* st);

static v3 hit_normal(const oscene* s, const ohit* h)
{
    int32_t t = h->prim;
    if (t >= s->ntris) return h->n; /* disk or sphere emitter */
    v3 n;
    if (!s->nidx || s->nidx[3 * t] < 0 || s->nidx[3 * t + 1] < 0 || s->nidx[3 * t + 2] < 0) {
        n = orc_normalize(vcross(s->e1[t], s->e2[t]));
    } else {
        v3 n0 = s->normals[s->nidx[3 * t]], n1 = s->normals[s->nidx[3 * t + 1]], n2 = s->normals[s->nidx[3 * t + 2]];
        /* normalize(n1*beta + n2*gamma + n0*(1-beta-gamma)) (CU/triangle_mesh.cu:69) */
        n = orc_normalize(vadd(vadd(vscale(n1, h->b), vscale(n2, h->g)), vscale(n0, (R)1 - h->b - h->g)));
    }
    /* normalize(rtTransformNormal(RT_OBJECT_TO_WORLD, shading_normal)): world-space geometry */
    return orc_normalize(n);
}

/* Differentiable form (new): nearest texel of a planar (x,z) map over
 * [uvp0,uvp2] x [uvp1,uvp3] (mode 0) or a spherical map about uvp[0..2] (mode 1).
 * No reference counterpart; pinned by texel-centre points of the rectangle and
 * of the sphere's (theta, phi) grid (tests/test_oracle_diff.py). */
static int diff_texel(int mode, const R* uvp, v3 p, int tw, int th)
{
    R u, v;
    if (mode == 0) {
        u = (p.x - uvp[0]) / (uvp[2] - uvp[0]);
        v = (p.z - uvp[1]) / (uvp[3] - uvp[1]);
    } else {
        v3 q = orc_normalize(V3(p.x - uvp[0], p.y - uvp[1], p.z - uvp[2]));
        u = (F_ATAN2(q.x, q.z) + PI_R) * (R)0.5f * INV_PI_R;
        v = F_ACOS(fmin((R)1, fmax((R)-1, q.y))) * INV_PI_R;
    }
    int iu = (int)floor(u * (R)tw), iv = (int)floor(v * (R)th);
    iu = iu < 0 ? 0 : (iu > tw - 1 ? tw - 1 : iu);
    iv = iv < 0 ? 0 : (iv > th - 1 ? th - 1 : iv);
    return iv * tw + iu;
}

void orc_diff_texel_batch(int n, int mode, const double* uvp, const double* pts, int tw, int th, int32_t* out)
{
    R u4[4] = {(R)uvp[0], (R)uvp[1], (R)uvp[2], (R)uvp[3]};
    for (int i = 0; i < n; i++)
        out[i] = diff_texel(mode, u4, V3((R)pts[3 * i], (R)pts[3 * i + 1], (R)pts[3 * i + 2]), tw, th);
}

/* CU/diffuse_shader.cu:42-107 (+ environment MIS, new). */
static v3 shade_lambert(const oscene* s, const orc_material* m, v3 o, v3 d, const ohit* h, int depth,
                        int* emit_light, uint32_t* t, ostats* st)
{
    v3 result = V3(0, 0, 0);
    if (depth > s->max_depth) return result;
    v3 hit_pos = vadd(o, vscale(d, h->t));
    v3 normal = hit_normal(s, h);
    v3 ffn = orc_faceforward(normal, vneg(d), normal);
    v3 rho = V3(m->diffuse[0], m->diffuse[1], m->diffuse[2]);
    R prob = (rho.x + rho.y + rho.z) / (R)3.0f;
    if (s->diff_on && s->alb_eval && (int32_t)(m - s->mats) == s->alb_mat) {
        /* textured albedo (new): evaluate at rho_eval, roulette on the frozen albedo */
        int k = diff_texel(s->alb_uv, s->alb_uvp, hit_pos, s->tex_w, s->tex_h);
        rho = V3(s->alb_eval[3 * k], s->alb_eval[3 * k + 1], s->alb_eval[3 * k + 2]);
        const R* f = s->alb_frozen + 3 * k;
        prob = (f[0] + f[1] + f[2]) / (R)3.0f;
    }
    uint32_t light_idx = (uint32_t)((R)s->nlights * orc_rnd_tea(t));
    if (s->nlights > 0) {
        if (light_idx >= (uint32_t)s->nlights) light_idx = (uint32_t)s->nlights - 1; /* new: see header */
        R light_pdf = (R)1 / (R)s->nlights;
        v3 w_l, rad;
        R dist, penv;
        direct_light(s, &s->lights[light_idx], hit_pos, t, &w_l, &rad, &dist, &penv);
        rad = V3(rad.x / light_pdf, rad.y / light_pdf, rad.z / light_pdf);
        R cos_theta = vdot(ffn, w_l);
        if (cos_theta > 0) {
            ohit sh;
            st->shd++;
            R V = trace(s, hit_pos, w_l, s->eps, dist - s->eps, 1, &sh) ? (R)0 : (R)1;
            v3 Li = vscale(rad, V);
            v3 c = vmul(vscale(vscale(Li, INV_PI_R), cos_theta), rho);
            if (penv >= 0) c = vscale(c, mis_power(penv * light_pdf, cos_theta * INV_PI_R));
            result = vadd(result, c);
            *emit_light = 0;
        }
    }
    R xi = orc_rnd_tea(t);
    if (xi < prob) {
        v3 nd = orc_sample_cosine_weighted(ffn, t);
        R pdf = s->has_env && s->env_is ? vdot(ffn, nd) * INV_PI_R : (R)-1;
        v3 child = radiance(s, hit_pos, nd, depth + 1, *emit_light, t, pdf, st);
        v3 cr = vmul(child, rho);
        result = vadd(result, V3(cr.x / prob, cr.y / prob, cr.z / prob));
    }
    return result;
}

/* CU/metallic_shader.cu:45-162, visible-normal branch. */
static v3 shade_conductor(const oscene* s, const orc_material* m, const cplx3* eta, v3 o, v3 d, const ohit* h,
                          int depth, uint32_t* seed, ostats* st)
{
    if (depth >= s->max_depth) return V3(0, 0, 0);
    v3 hit_point = vadd(o, vscale(d, h->t));
    v3 normal = hit_normal(s, h);
    v3 ffn = orc_faceforward(normal, vneg(d), normal);
    v3 w_i = vneg(d);
    R ax = m->roughness[0], ay = m->roughness[1];
    R ae = ax;  /* evaluation roughness of the differentiable form */
    if (s->diff_on && s->rough_eval && (int32_t)(m - s->mats) == s->rough_mat) {
        /* textured roughness (new): sample with the frozen alpha, evaluate at alpha_eval */
        int k = diff_texel(s->rough_uv, s->rough_uvp, hit_point, s->tex_w, s->tex_h);
        ax = ay = s->rough_frozen[k];
        ae = s->rough_eval[k];
    }
    if (m->microfacet_model == 2) {
        /* multiple-scattering model (CU/metallic_shader.cu:66-82): the GGX walk in
         * the frame of the unflipped normal, child radiance times its weight */
        v3 lo, wt;
        int order;
        const cplx3* ie = s->ieta + (eta - s->eta);
        int ok = m->normal_distribution == 1
                     ? orc_ggx_ms_conductor_sample(orc_to_local(w_i, normal), eta, ie, ax, ay, seed, MS_WALK_CAP, &lo,
                                                   &wt, &order)
                     : orc_beckmann_ms_conductor_sample(orc_to_local(w_i, normal), eta, ie, ax, ay, seed, MS_WALK_CAP,
                                                        &lo, &wt, &order);
        if (!ok) return V3(0, 0, 0);
        v3 child = radiance(s, hit_point, orc_to_world(lo, normal), depth + 1, 1, seed, (R)-1, st);
        return vmul(child, wt);
    }
    R z1 = orc_rnd_tea(seed);
    R z2 = orc_rnd_tea(seed);
    int dist = m->normal_distribution;
    if (m->microfacet_model == 0) {
        /* Walter model (CU/metallic_shader.cu:92-148): normal ~ D cos, weight
         * F |i.m| G1(i) G1(o) / (|i.n| |n.m|) */
        v3 mw = orc_sample_normal_walter(ffn, z1, z2, ax, dist);
        R gi = orc_masking_iso(w_i, mw, ffn, ax, dist);
        if (gi <= 0) return V3(0, 0, 0);
        v3 Fw = orc_fresnel_conductor(w_i, mw, eta);
        v3 wo = vadd(vneg(w_i), vscale(vscale(mw, (R)2.0f), vdot(w_i, mw)));
        R go = orc_masking_iso(wo, mw, ffn, ax, dist);
        if (go <= 0) return V3(0, 0, 0);
        v3 cw = radiance(s, hit_point, wo, depth + 1, 1, seed, (R)-1, st);
        R w = fabs(vdot(w_i, mw)) * gi * go / (fabs(vdot(w_i, ffn)) * fabs(vdot(normal, mw)));
        return vmul(cw, vscale(Fw, w));
    }
    if (dist == 0) {
        /* Beckmann visible normals (CU/metallic_shader.cu:96-146) */
        v3 mb = orc_to_world(orc_beckmann_sample_vndf(orc_to_local(w_i, ffn), ax, ay, z1, z2), ffn);
        R gi = orc_beckmann_G1(w_i, mb, ffn, ax, ay);
        if (gi <= 0) return V3(0, 0, 0);
        v3 Fb = orc_fresnel_conductor(w_i, mb, eta);
        v3 wo = vadd(vneg(w_i), vscale(vscale(mb, (R)2.0f), vdot(w_i, mb)));
        R go = orc_beckmann_G1(wo, mb, ffn, ax, ay);
        if (go <= 0) return V3(0, 0, 0);
        v3 cb = radiance(s, hit_point, wo, depth + 1, 1, seed, (R)-1, st);
        return vmul(cb, vscale(Fb, go));
    }
    v3 mn = orc_to_world(orc_ggx_sample_vndf(orc_to_local(w_i, ffn), ax, ay, z1, z2), ffn);
    R G_i = orc_ggx_G1(w_i, mn, ffn, ax, ay);
    if (G_i <= 0) return V3(0, 0, 0);
    v3 F = orc_fresnel_conductor(w_i, mn, eta);
    v3 w_o = vadd(vneg(w_i), vscale(vscale(mn, (R)2.0f), vdot(w_i, mn)));
    R G_o = orc_ggx_G1(w_o, mn, ffn, ax, ay);
    if (G_o <= 0) return V3(0, 0, 0);
    v3 child = radiance(s, hit_point, w_o, depth + 1, 1, seed, (R)-1, st);
    v3 weight = vscale(F, G_o);
    if (ae != ax) {
        /* f cos / pdf at alpha_eval for a direction drawn with the frozen alpha:
         * F G1(wi) G1(wo) D(m) / (G1_0(wi) D_0(m)) with the separable G2 the
         * reference uses (CU/metallic_shader.cu:117-121) */
        v3 ml = orc_to_local(mn, ffn);
        R num = orc_ggx_G1(w_i, mn, ffn, ae, ae) * orc_ggx_G1(w_o, mn, ffn, ae, ae) * orc_ggx_eval_ndf(ml, ae, ae);
        R den = G_i * orc_ggx_eval_ndf(ml, ax, ay);
        weight = vscale(F, num / den);
    }
    return vmul(child, weight);
}

/* masking_G1(v, m, n, a_x, a_y, distribution) (ORF/Microfacet.h:24-32). */
static R masking_xy(v3 v, v3 m, v3 n, R ax, R ay, int dist)
{
    return dist == 1 ? orc_ggx_G1(v, m, n, ax, ay) : orc_beckmann_G1(v, m, n, ax, ay);
}

/* masking_G1 / microfacet_sample_visible_normal dispatch (ORF/Microfacet.h:24-32, 72-83). */
static R masking(v3 v, v3 m, v3 n, R a, int dist)
{
    return dist == 1 ? orc_ggx_G1(v, m, n, a, a) : orc_beckmann_G1(v, m, n, a, a);
}

static v3 sample_visible(v3 w_i, v3 n, R a, R z1, R z2, int dist)
{
    v3 l = orc_to_local(w_i, n);
    v3 m = dist == 1 ? orc_ggx_sample_vndf(l, a, a, z1, z2) : orc_beckmann_sample_vndf(l, a, a, z1, z2);
    return orc_to_world(m, n);
}

/* CU/rough_transparent_shader.cu:66-225, visible-normal model, with
 * get_rough_glass_rays (:39-55). */
static v3 shade_dielectric(const oscene* s, const orc_material* m, v3 o, v3 d, const ohit* h, int depth,
                           uint32_t* t, ostats* st)
{
    v3 zero = V3(0, 0, 0);
    if (depth >= s->max_depth) return zero;
    v3 hit_point = vadd(o, vscale(d, h->t));
    v3 normal = hit_normal(s, h);
    v3 ffn = orc_faceforward(normal, vneg(d), normal);
    v3 w_i = vneg(d);
    v3 beam = V3(1, 1, 1);
    R cos_theta = vdot(vneg(d), normal);
    if (cos_theta < 0) {
        beam = V3(F_EXP(-h->t * (R)m->absorption[0]), F_EXP(-h->t * (R)m->absorption[1]), F_EXP(-h->t * (R)m->absorption[2]));
        R prob = (beam.x + beam.y + beam.z) / (R)3.0f;
        if (orc_rnd_tea(t) >= prob) return zero;
        beam = V3(beam.x / prob, beam.y / prob, beam.z / prob);
    }
    R z1 = orc_rnd_tea(t);
    R z2 = orc_rnd_tea(t);
    R a = m->roughness[0];
    int dist = m->normal_distribution;
    if (m->microfacet_model == 2) {
        /* multiple-scattering model (CU/rough_transparent_shader.cu:98-116): z1, z2
         * above are drawn and unused; the walk runs in the frame of `normal` */
        v3 lo;
        int order;
        int ok = dist == 1 ? orc_ggx_ms_dielectric_sample(orc_to_local(w_i, normal), (R)1.0f / (R)m->ior, a, a, t,
                                                          MS_WALK_CAP, &lo, &order)
                           : orc_beckmann_ms_dielectric_sample(orc_to_local(w_i, normal), (R)1.0f / (R)m->ior, a, a, t,
                                                               MS_WALK_CAP, &lo, &order);
        if (!ok) return zero;
        v3 dir = orc_to_world(lo, normal);
        v3 child = radiance(s, hit_point, dir, depth + 1, 1, t, (R)-1, st);
        return vmul(child, beam);
    }
    int walter = m->microfacet_model == 0;
    v3 mn = walter ? orc_sample_normal_walter(ffn, z1, z2, a, dist) : sample_visible(w_i, ffn, a, z1, z2, dist);
    R eta = vdot(normal, w_i) < 0 ? (R)m->ior : (R)1.0f / (R)m->ior;
    R G_i = walter ? orc_masking_iso(w_i, mn, ffn, a, dist) : masking(w_i, mn, ffn, a, dist);
    if (G_i <= 0) return zero;
    R ct = fabs(vdot(mn, vneg(d)));
    R st2 = eta * eta * ((R)1.0f - ct * ct);
    R ctt = sqrt((R)1.0f - st2);
    R Rf = st2 < (R)1.0f ? orc_fresnel_R3(ct, ctt, eta) : (R)1.0f;
    v3 refr = vadd(vscale(d, eta), vscale(mn, eta * ct - ctt));
    v3 refl = vsub(d, vscale(vscale(mn, (R)2.0f), vdot(mn, d)));
    R rr = orc_rnd_tea(t);
    v3 dir = rr > Rf ? refr : refl;
    R G_o = walter ? orc_masking_iso(dir, mn, ffn, a, dist) : masking(dir, mn, ffn, a, dist);
    if (G_o <= 0) return zero;
    v3 child = radiance(s, hit_point, dir, depth + 1, 1, t, (R)-1, st);
    /* Walter weight |i.m| G1(i) G1(o) / (|i.n| |n.m|) (CU/rough_transparent_shader.cu:184-200) */
    R w = walter ? fabs(vdot(w_i, mn)) * G_i * G_o / (fabs(vdot(w_i, ffn)) * fabs(vdot(normal, mn))) : G_o;
    v3 result = vscale(child, w);
    return vmul(result, beam);
}

/* Fresnel reflectance at incidence cosine c through relative index eta; 1 under
 * total internal reflection (CU/lambertian_interface_shader.cu:68-75). */
static R two_C1(R n);

/* CU/rough_diffuse_shader.cu:40-290: a Lambert base under a rough dielectric.
 * NEE (a light other than the environment, shadow ray to the full light
 * distance) adds Li/pi cos_l rho n12^2 T_i T_d / (1 - rho r10) G1(l, m) and
 * Li R_r G1(i, h) G1(l, h) D_wi(h) / (4 |i.n l.n|), and turns emission off once
 * a light is seen; the continuation is the cosine bounce about the sampled
 * microfacet normal m (weight rho n12^2 T_r / (1 - rho r10) G1(o, m)) when
 * xi > R_i, else the mirror about m (weight G1(o, m)).  The multiple-scattering
 * model (:126-139, :187-216, :252-268) takes Li (diffuse walk estimate +
 * interface walk estimate with the interface as a conductor of index n12),
 * drawn before the shadow ray, and continues with the diffuse walk (weight
 * rho^k n12^2 T / (1 - rho r10), T at its last visible normal) or the
 * conductor walk in the un-flipped normal's frame (weight 1). */
static v3 shade_rough_diffuse(const oscene* s, const orc_material* m, v3 o, v3 d, const ohit* h, int depth,
                              int emit_light, uint32_t* t, ostats* st)
{
    v3 result = V3(0, 0, 0);
    if (depth > s->max_depth) return result;
    v3 hit_pos = vadd(o, vscale(d, h->t));
    v3 normal = hit_normal(s, h);
    v3 ffn = orc_faceforward(normal, vneg(d), normal);
    v3 w_i = vneg(d);
    R n12 = (R)1.0f / (R)m->ior;
    R r10 = two_C1((R)m->ior);
    R ax = m->roughness[0], ay = m->roughness[1];
    int nd = m->normal_distribution;
    v3 rho = V3(m->diffuse[0], m->diffuse[1], m->diffuse[2]);
    v3 denom = V3((R)1.0f - rho.x * r10, (R)1.0f - rho.y * r10, (R)1.0f - rho.z * r10);
    int ms = m->microfacet_model == 2;
    cplx e12 = {n12, 0};
    cplx ie12 = orc_cdiv_fc((R)1.0f, e12);
    cplx3 eta3 = {e12, e12, e12}, ieta3 = {ie12, ie12, ie12};
    R z1 = orc_rnd_tea(t);
    R z2 = orc_rnd_tea(t);
    v3 l = orc_to_local(w_i, ffn);
    v3 mn = orc_to_world(nd == 1 ? orc_ggx_sample_vndf(l, ax, ay, z1, z2) : orc_beckmann_sample_vndf(l, ax, ay, z1, z2),
                         ffn);
    R ci = vdot(w_i, mn);
    R Ri = 1;
    R st2 = n12 * n12 * ((R)1.0f - ci * ci);
    if (st2 < (R)1.0f) {
        R ct = sqrt((R)1.0f - st2);
        Ri = orc_fresnel_R3(ci, ct, n12);
    }
    R Ti = (R)1.0f - Ri;
    uint32_t light_idx = (uint32_t)((R)s->nlights * orc_rnd_tea(t));
    if (s->nlights > 0) {
        if (light_idx >= (uint32_t)s->nlights) light_idx = (uint32_t)s->nlights - 1;
        R light_pdf = (R)1 / (R)s->nlights;
        v3 w_l, rad;
        R dist, penv;
        direct_light(s, &s->lights[light_idx], hit_pos, t, &w_l, &rad, &dist, &penv);
        rad = V3(rad.x / light_pdf, rad.y / light_pdf, rad.z / light_pdf);
        R cl = vdot(ffn, w_l);
        if (cl > 0 && penv < 0) {
            v3 msc = V3(0, 0, 0);
            if (ms) {
                v3 lwi = orc_to_local(w_i, ffn), lwl = orc_to_local(w_l, ffn);
                v3 fd = orc_ms_diffuse_eval(nd, lwi, lwl, ax, ay, rho, t, MS_DIFFUSE_CAP);
                v3 fr = orc_ms_conductor_eval(nd, lwi, lwl, &eta3, &ieta3, ax, ay, t, MS_WALK_CAP);
                msc = vmul(rad, vadd(fd, fr));
            }
            ohit sh;
            st->shd++;
            if (ms) {
                if (!trace(s, hit_pos, w_l, s->eps, dist, 1, &sh)) {
                    result = vadd(result, msc);
                    emit_light = 0;
                }
            } else if (!trace(s, hit_pos, w_l, s->eps, dist, 1, &sh)) {
                v3 rn = orc_normalize(vadd(w_i, w_l));
                R cr = vdot(rn, w_l);
                R Rr = 1;
                st2 = n12 * n12 * ((R)1.0f - cr * cr);
                if (st2 < (R)1.0f) {
                    R ct = sqrt((R)1.0f - st2);
                    Rr = orc_fresnel_R3(cr, ct, n12);
                }
                R cd = vdot(mn, w_l);
                R Rd = 1;
                st2 = n12 * n12 * ((R)1.0f - cd * cd);
                if (st2 < (R)1.0f) {
                    R ct = sqrt((R)1.0f - st2);
                    Rd = orc_fresnel_R3(cd, ct, n12);
                }
                R Td = (R)1.0f - Rd;
                R gi = masking_xy(w_i, rn, ffn, ax, ay, nd);
                R go = masking_xy(w_l, rn, ffn, ax, ay, nd);
                R god = masking_xy(w_l, mn, ffn, ax, ay, nd);
                R D = orc_eval_visible_normal(w_i, rn, ffn, ax, ay, nd);
                R wr = gi * go * D / ((R)4.0f * fabs(vdot(w_i, ffn) * vdot(w_l, ffn)));
                v3 a = vscale(vscale(rad, INV_PI_R), cl);
                a = vmul(a, rho);
                a = vscale(vscale(vscale(vscale(a, n12), n12), Ti), Td);
                a = V3(a.x / denom.x, a.y / denom.y, a.z / denom.z);
                a = vscale(a, god);
                v3 b = vscale(vscale(rad, Rr), wr);
                result = vadd(result, vadd(a, b));
                emit_light = 0;
            }
        }
    }
    R xi = orc_rnd_tea(t);
    if (ms) {
        v3 lo, k;
        int order;
        if (xi > Ri) {
            v3 lm;
            orc_ms_diffuse_sample(nd, orc_to_local(w_i, ffn), ax, ay, rho, t, MS_DIFFUSE_CAP, &lo, &lm, &k, &order);
            v3 wo = orc_to_world(lo, ffn);
            R cr = vdot(wo, orc_to_world(lm, ffn));
            R Rr = 1;
            st2 = n12 * n12 * ((R)1.0f - cr * cr);
            if (st2 < (R)1.0f) Rr = orc_fresnel_R3(cr, sqrt((R)1.0f - st2), n12);
            v3 child = radiance(s, hit_pos, wo, depth + 1, emit_light, t, (R)-1, st);
            v3 c = vscale(vscale(vscale(child, n12), n12), (R)1.0f - Rr);
            c = V3(c.x / denom.x, c.y / denom.y, c.z / denom.z);
            result = vadd(result, vmul(c, k));
        } else {
            v3 lwi = orc_to_local(w_i, normal);
            int ok = nd == 1 ? orc_ggx_ms_conductor_sample(lwi, &eta3, &ieta3, ax, ay, t, MS_WALK_CAP, &lo, &k, &order)
                             : orc_beckmann_ms_conductor_sample(lwi, &eta3, &ieta3, ax, ay, t, MS_WALK_CAP, &lo, &k,
                                                                &order);
            if (!ok) return result;
            result = vadd(result, radiance(s, hit_pos, orc_to_world(lo, normal), depth + 1, emit_light, t, (R)-1, st));
        }
        return result;
    }
    if (xi > Ri) {
        v3 dd = orc_sample_cosine_weighted(mn, t);
        R cr = vdot(dd, mn);
        R gom = masking_xy(dd, mn, ffn, ax, ay, nd);
        if (gom < 0) return result;
        R Rr = 1;
        st2 = n12 * n12 * ((R)1.0f - cr * cr);
        if (st2 < (R)1.0f) {
            R ct = sqrt((R)1.0f - st2);
            Rr = orc_fresnel_R3(cr, ct, n12);
        }
        v3 child = radiance(s, hit_pos, dd, depth + 1, emit_light, t, (R)-1, st);
        v3 c = vscale(vscale(vscale(vmul(child, rho), n12), n12), (R)1.0f - Rr);
        c = V3(c.x / denom.x, c.y / denom.y, c.z / denom.z);
        result = vadd(result, vscale(c, gom));
    } else {
        v3 rd = vsub(d, vscale(vscale(mn, (R)2.0f), vdot(mn, d)));
        R gom = masking_xy(rd, mn, ffn, ax, ay, nd);
        if (gom < 0) return result;
        v3 child = radiance(s, hit_pos, rd, depth + 1, emit_light, t, (R)-1, st);
        result = vadd(result, vscale(child, gom));
    }
    return result;
}

/* CU/flat_shader.cu:40-127: the flat colour, times the picked light's
 * visibility with flat_shadow; with highlight, a visible light whose microfacet
 * reflection Le G1(i) G1(o) D_wi R / (4 |i.n l.n|) has luminance above the
 * threshold shows the highlight colour.  The microfacet normal the shader
 * samples is never used: only its two random numbers are drawn. */
static v3 shade_flat(const oscene* s, const orc_material* m, v3 o, v3 d, const ohit* h, int depth, uint32_t* t,
                     ostats* st)
{
    if (depth > s->max_depth) return V3(0, 0, 0);
    v3 hit_pos = vadd(o, vscale(d, h->t));
    v3 normal = hit_normal(s, h);
    v3 ffn = orc_faceforward(normal, vneg(d), normal);
    v3 w_i = vneg(d);
    v3 result = V3(m->diffuse[0], m->diffuse[1], m->diffuse[2]);
    orc_rnd_tea(t);
    orc_rnd_tea(t);
    R V = 0;
    uint32_t light_idx = (uint32_t)((R)s->nlights * orc_rnd_tea(t));
    v3 w_l = V3(0, 0, 0), rad = V3(0, 0, 0);
    if (s->nlights > 0) {
        if (light_idx >= (uint32_t)s->nlights) light_idx = (uint32_t)s->nlights - 1;
        R light_pdf = (R)1 / (R)s->nlights;
        R dist, penv;
        direct_light(s, &s->lights[light_idx], hit_pos, t, &w_l, &rad, &dist, &penv);
        rad = V3(rad.x / light_pdf, rad.y / light_pdf, rad.z / light_pdf);
        if (vdot(ffn, w_l) > 0) {
            ohit sh;
            st->shd++;
            V = trace(s, hit_pos, w_l, s->eps, dist - s->eps, 1, &sh) ? (R)0 : (R)1;
        }
    }
    if (m->flat_flags & 1) result = vscale(result, V);
    if ((m->flat_flags & 2) && V > 0) {
        v3 rn = orc_normalize(vadd(w_i, w_l));
        R cr = vdot(rn, w_l), n12 = (R)1.0f / (R)m->ior;
        R Rf = 1;
        R st2 = n12 * n12 * ((R)1.0f - cr * cr);
        if (st2 < (R)1.0f) {
            R ct = sqrt((R)1.0f - st2);
            Rf = orc_fresnel_R3(cr, ct, n12);
        }
        R ax = m->roughness[0], ay = m->roughness[1];
        int nd = m->normal_distribution;
        R gi = masking_xy(w_i, rn, ffn, ax, ay, nd);
        R go = masking_xy(w_l, rn, ffn, ax, ay, nd);
        R D = orc_eval_visible_normal(w_i, rn, ffn, ax, ay, nd);
        R wr = gi * go * D * Rf / ((R)4.0f * fabs(vdot(w_i, ffn) * vdot(w_l, ffn)));
        v3 spec = vscale(rad, wr);
        R lum = (R)0.2126f * spec.x + (R)0.7152f * spec.y + (R)0.0722f * spec.z;
        if (lum > (R)m->flat_threshold) result = vscale(V3(m->radiance[0], m->radiance[1], m->radiance[2]), V);
    }
    return result;
}

/* CU/transparent_shader.cu:51-114 (DIFFUSE_PART off): depth > max_depth ends the
 * path; inside, Beer absorption with Russian roulette on its mean and the
 * normal and index flipped; the Fresnel reflectance picks mirror or refraction. */
static v3 shade_transparent(const oscene* s, const orc_material* m, v3 o, v3 d, const ohit* h, int depth,
                            uint32_t* t, ostats* st)
{
    v3 zero = V3(0, 0, 0);
    if (depth > s->max_depth) return zero;
    v3 hit_point = vadd(o, vscale(d, h->t));
    v3 normal = hit_normal(s, h);
    v3 beam = V3(1, 1, 1);
    R n12 = (R)1.0f / (R)m->ior;
    R c = vdot(vneg(d), normal);
    if (c < 0) {
        beam = V3(F_EXP(-h->t * (R)m->absorption[0]), F_EXP(-h->t * (R)m->absorption[1]),
                  F_EXP(-h->t * (R)m->absorption[2]));
        R prob = (beam.x + beam.y + beam.z) / (R)3.0f;
        if (orc_rnd_tea(t) >= prob) return zero;
        beam = V3(beam.x / prob, beam.y / prob, beam.z / prob);
        n12 = (R)m->ior;
        normal = vneg(normal);
        c = -c;
    }
    R Rf = 1, ct = 0;
    R st2 = n12 * n12 * ((R)1.0f - c * c);
    if (st2 < (R)1.0f) {
        ct = sqrt((R)1.0f - st2);
        Rf = orc_fresnel_R3(c, ct, n12);
    }
    R xi = orc_rnd_tea(t);
    v3 dir = xi < Rf ? vsub(d, vscale(vscale(normal, (R)2.0f), vdot(normal, d)))
                     : vadd(vscale(d, n12), vscale(normal, n12 * c - ct));
    v3 child = radiance(s, hit_point, dir, depth + 1, 1, t, (R)-1, st);
    return vmul(child, beam);
}

static R fresnel_R_cos(R c, R eta)
{
    R st2 = eta * eta * ((R)1.0f - c * c);
    if (!(st2 < (R)1.0f)) return 1;
    return orc_fresnel_R3(c, sqrt((R)1.0f - st2), eta);
}

/* two_C1 (ORF/Fresnel.h:93-101), the published polynomial fits. */
static R two_C1(R n)
{
    if (n >= (R)1.0f)
        return (R)-9.23372f + n * ((R)22.2272f + n * ((R)-20.9292f + n * ((R)10.2291f + n * ((R)-2.54396f + n * (R)0.254913f))));
    return (R)0.919317f + n * ((R)-3.4793f + n * ((R)6.75335f + n * ((R)-7.80989f + n * ((R)4.98554f - n * (R)1.36881f))));
}

/* CU/lambertian_interface_shader.cu:51-173 (+ environment MIS, new). */
static v3 shade_interface(const oscene* s, const orc_material* m, v3 o, v3 d, const ohit* h, int depth,
                          int emit_light, uint32_t* t, ostats* st)
{
    v3 result = V3(0, 0, 0);
    if (depth > s->max_depth) return result;
    v3 hit_pos = vadd(o, vscale(d, h->t));
    v3 normal = hit_normal(s, h);
    v3 ffn = orc_faceforward(normal, vneg(d), normal);
    v3 rho = V3(m->diffuse[0], m->diffuse[1], m->diffuse[2]);
    R ior = (R)m->ior;
    R n12 = (R)1.0f / ior;
    R r10 = two_C1(ior);
    v3 den = V3((R)1.0f - rho.x * r10, (R)1.0f - rho.y * r10, (R)1.0f - rho.z * r10);
    R cos_i = vdot(vneg(d), ffn);
    R R_i = fresnel_R_cos(cos_i, n12);
    R T_i = (R)1.0f - R_i;
    uint32_t light_idx = (uint32_t)((R)s->nlights * orc_rnd_tea(t));
    if (s->nlights > 0) {
        if (light_idx >= (uint32_t)s->nlights) light_idx = (uint32_t)s->nlights - 1; /* new: see header */
        R light_pdf = (R)1 / (R)s->nlights;
        v3 w_l, rad;
        R dist, penv;
        direct_light(s, &s->lights[light_idx], hit_pos, t, &w_l, &rad, &dist, &penv);
        rad = V3(rad.x / light_pdf, rad.y / light_pdf, rad.z / light_pdf);
        R cos_l = vdot(ffn, w_l);
        if (cos_l > 0) {
            ohit sh;
            st->shd++;
            R V = trace(s, hit_pos, w_l, s->eps, dist, 1, &sh) ? (R)0 : (R)1;  /* tmax = dist (:100) */
            v3 Li = vscale(rad, V);
            R T_l = (R)1.0f - fresnel_R_cos(cos_l, n12);
            v3 c = vmul(vscale(vscale(Li, INV_PI_R), cos_l), rho);
            c = vscale(vscale(vscale(vscale(c, n12), n12), T_l), T_i);
            c = V3(c.x / den.x, c.y / den.y, c.z / den.z);
            if (penv >= 0) c = vscale(c, mis_power(penv * light_pdf, cos_l * INV_PI_R));
            result = vadd(result, c);
            emit_light = 0;
        }
    }
    R xi = orc_rnd_tea(t);
    if (xi > R_i) {
        v3 nd = orc_sample_cosine_weighted(ffn, t);
        R cos_r = vdot(nd, ffn);
        R T_r = (R)1.0f - fresnel_R_cos(cos_r, n12);
        R pdf = s->has_env && s->env_is ? cos_r * INV_PI_R : (R)-1;
        v3 child = radiance(s, hit_pos, nd, depth + 1, emit_light, t, pdf, st);
        v3 cr = vscale(vscale(vscale(vmul(child, rho), n12), n12), T_r);
        result = vadd(result, V3(cr.x / den.x, cr.y / den.y, cr.z / den.z));
    } else {
        /* reflect(ray.direction, normal): i - 2 n dot(n, i) */
        v3 rd = vsub(d, vscale(vscale(normal, (R)2.0f), vdot(normal, d)));
        v3 child = radiance(s, hit_pos, rd, depth + 1, 1, t, (R)-1, st);
        result = vadd(result, child);
    }
    return result;
}

/* ---------------------------------------------------------------- translucent (new estimator)
 * Coefficients: ORF/RoughTranslucentMaterial.cpp:366-395 (outside ior 1); Fresnel
 * moments ORF/Fresnel.h:93-124; dipoles ORF/dipoles/standard_dipole.h:8-29 and
 * ORF/dipoles/directional_dipole.h:8-72; shader structure
 * CU/subsurface_scattering_shader.cu:55-198; probe sample CU/sample_camera.cu:109-212
 * (TEA streams instead of RND_64).  The gather replaces the reference's sum over
 * all probes by exit-point importance sampling (probe rays along -n through
 * exponentially distributed tangent-plane points, nearest probe within `cell`),
 * the estimator the product uses. */

static v3 vdiv(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static v3 vsqrt3(v3 a) { return V3(sqrt(a.x), sqrt(a.y), sqrt(a.z)); }
static v3 vexp3(v3 a) { return V3(F_EXP(a.x), F_EXP(a.y), F_EXP(a.z)); }
static v3 vaddf(v3 a, R f) { return V3(a.x + f, a.y + f, a.z + f); }
static v3 vfsub(R f, v3 a) { return V3(f - a.x, f - a.y, f - a.z); }

static R three_C2(R n)
{
    R r;
    if (n >= (R)1.0f) {
        r = (R)-1641.1f + n * ((R)1213.67f + n * ((R)-568.556f + n * ((R)164.798f + n * ((R)-27.0181f + n * (R)1.91826f))));
        r += ((((R)135.926f / n) - (R)656.175f) / n + (R)1376.53f) / n;
    } else {
        r = (R)0.828421f + n * ((R)-2.62051f + n * ((R)3.36231f + n * ((R)-1.95284f + n * ((R)0.236494f + n * (R)0.145787f))));
    }
    return r;
}
static R C_phi(R ni) { return (R)0.25f * ((R)1.0f - two_C1(ni)); }
static R C_E(R ni) { return (R)0.5f * ((R)1.0f - three_C2(ni)); }

/* Coefficients of material m with its absorption and scattering (1/mm) given
 * as sa[3], ss[3] (the differentiable form's parameters). */
static osss sss_coefficients_with(const orc_material* m, const double* sa, const double* ss)
{
    osss p;
    memset(&p, 0, sizeof(p));
    R scale = (R)m->sss_scale;
    p.dipole_model = m->dipole_model;
    p.absorption = vscale(V3((R)sa[0], (R)sa[1], (R)sa[2]), scale);
    p.scattering = vscale(V3((R)ss[0], (R)ss[1], (R)ss[2]), scale);
    p.meancosine = V3(m->sss_meancosine[0], m->sss_meancosine[1], m->sss_meancosine[2]);
    p.relative_ior = (R)m->ior;
    R ior = (R)m->ior;
    v3 reducedScattering = vmul(p.scattering, vfsub((R)1.0f, p.meancosine));
    p.relative_ior /= ((R)1.0f + (R)1.0f + (R)1.0f) / (R)3.0f;
    p.extinction = vadd(p.scattering, p.absorption);
    p.reducedExtinction = vadd(reducedScattering, p.absorption);
    p.deltaEddExtinction = vadd(vmul(p.scattering, vfsub((R)1.0f, vmul(p.meancosine, p.meancosine))), p.absorption);
    p.D = vdiv(V3(1, 1, 1), vscale(p.reducedExtinction, (R)3.0f));
    p.transport = vsqrt3(vdiv(p.absorption, p.D));
    p.C_phi = V3(C_phi(ior), C_phi(ior), C_phi(ior));
    p.C_phi_inv = V3(C_phi((R)1.0f / ior), C_phi((R)1.0f / ior), C_phi((R)1.0f / ior));
    p.C_E = V3(C_E(ior), C_E(ior), C_E(ior));
    p.albedo = vdiv(p.scattering, p.extinction);
    p.reducedAlbedo = vdiv(reducedScattering, p.reducedExtinction);
    p.de = vdiv(vscale(p.D, (R)2.131f), vsqrt3(p.reducedAlbedo));
    p.A = vdiv(vfsub((R)1.0f, p.C_E), vscale(p.C_phi, (R)2.0f));
    p.three_D = vscale(p.D, (R)3.0f);
    p.two_A_de = vmul(vscale(p.A, (R)2.0f), p.de);
    v3 g = vscale(vdiv(V3(1, 1, 1), vscale(p.C_phi_inv, (R)4.0f)), (R)1.0f);
    R q = (R)4.0f * PI_R * PI_R;
    p.global_coeff = V3(g.x / q, g.y / q, g.z / q);
    p.one_over_three_ext = vdiv(V3(1, 1, 1), vscale(p.extinction, (R)3.0f));
    p.mean_transport = (p.transport.x + p.transport.y + p.transport.z) / (R)3.0f;
    return p;
}

static osss sss_coefficients(const orc_material* m)
{
    double sa[3], ss[3];
    for (int k = 0; k < 3; k++) {
        sa[k] = (double)m->sss_absorption[k];
        ss[k] = (double)m->sss_scattering[k];
    }
    return sss_coefficients_with(m, sa, ss);
}

static v3 dipole_bssrdf(R dist, const osss* p)
{
    v3 real_source = vmul(p->three_D, p->three_D);
    v3 extrapolation = vmul(vscale(p->A, (R)4.0f), p->D);
    v3 virtual_source = vmul(extrapolation, extrapolation);
    v3 corrected_mean_free = vadd(p->three_D, extrapolation);
    R r2 = dist * dist;
    v3 r_sqr = V3(r2, r2, r2);
    v3 d_r_sqr = vadd(r_sqr, real_source);
    v3 d_r = vsqrt3(d_r_sqr);
    v3 d_v_sqr = vadd(r_sqr, virtual_source);
    v3 d_v = vsqrt3(d_v_sqr);
    v3 tr_r = vmul(p->transport, d_r);
    v3 S_r = vdiv(vmul(p->three_D, vaddf(tr_r, (R)1.0f)), vmul(d_r_sqr, d_r));
    S_r = vmul(S_r, vexp3(vneg(tr_r)));
    v3 tr_v = vmul(p->transport, d_v);
    v3 S_v = vdiv(vmul(corrected_mean_free, vaddf(tr_v, (R)1.0f)), vmul(d_v_sqr, d_v));
    S_v = vmul(S_v, vexp3(vneg(tr_v)));
    return vadd(S_r, S_v);
}

static v3 S_infinite(v3 r_sqr, v3 x_dot_w12, R no_dot_w12, v3 x_dot_no, const osss* p)
{
    v3 r = vsqrt3(r_sqr);
    v3 r_tr = vmul(p->transport, r);
    v3 r_tr_p1 = vaddf(r_tr, (R)1.0f);
    v3 T = vexp3(vneg(r_tr));
    v3 coeff = vdiv(T, vmul(r, r_sqr));
    v3 first = vscale(vmul(p->C_phi, vadd(vmul(r_sqr, p->reducedExtinction), vmul(r_tr_p1, x_dot_w12))), (R)3.0f);
    v3 inner = vadd(r_tr_p1, vmul(vdiv(vmul(p->three_D, vadd(vscale(r_tr_p1, (R)3.0f), vmul(r_tr, r_tr))), r_sqr), x_dot_w12));
    v3 second = vmul(p->C_E, vsub(vscale(vmul(p->three_D, r_tr_p1), no_dot_w12), vmul(inner, x_dot_no)));
    return vmul(coeff, vsub(first, second));
}

static v3 dirpole_bssrdf(v3 xi, v3 ni, v3 w12, v3 xo, v3 no, const osss* p)
{
    v3 x = vsub(xo, xi);
    R r_sqr = vdot(x, x);
    v3 dr_sqr = V3(r_sqr, r_sqr, r_sqr);
    R dot_x_w12 = vdot(x, w12);
    R num = r_sqr - dot_x_w12 * dot_x_w12;
    v3 cos_beta = vneg(vsqrt3(vdiv(V3(num, num, num), vaddf(vmul(p->de, p->de), r_sqr))));
    R mu0 = -vdot(no, w12);
    R edge = mu0 > 0 ? (R)1.0f : (R)0.0f;
    v3 D_prime = vadd(vscale(vscale(p->D, mu0), edge), vscale(p->one_over_three_ext, (R)1.0f - edge));
    dr_sqr = vadd(dr_sqr, vmul(D_prime, vsub(D_prime, vscale(vmul(vscale(p->de, (R)2.0f), cos_beta), edge))));
    v3 t = orc_normalize(vcross(ni, x));
    v3 nistar = r_sqr < (R)1.0e-12f ? ni : vcross(orc_normalize(x), t);
    v3 wv = vsub(w12, vscale(nistar, (R)2.0f * vdot(w12, nistar)));
    v3 xr = vsub(x, vscale(nistar, p->two_A_de.x)), xg = vsub(x, vscale(nistar, p->two_A_de.y)),
       xb = vsub(x, vscale(nistar, p->two_A_de.z));
    v3 dv_sqr = V3(vdot(xr, xr), vdot(xg, xg), vdot(xb, xb));
    v3 x_dot_wv = V3(vdot(xr, wv), vdot(xg, wv), vdot(xb, wv));
    v3 x_dot_no = V3(vdot(xr, no), vdot(xg, no), vdot(xb, no));
    R xn = vdot(x, no);
    v3 Sr = S_infinite(dr_sqr, V3(dot_x_w12, dot_x_w12, dot_x_w12), -mu0, V3(xn, xn, xn), p);
    v3 Sv = S_infinite(dv_sqr, x_dot_wv, vdot(no, wv), x_dot_no, p);
    v3 Sd = vsub(Sr, Sv);
    return V3(fmax(Sd.x, (R)0), fmax(Sd.y, (R)0), fmax(Sd.z, (R)0));
}

/* Nearest probe strictly within `cell` of y (ties: lower index), by brute force. */
static int nearest_probe(const oscene* s, v3 y)
{
    R best = s->probe_cell * s->probe_cell;
    int bi = -1;
    for (int j = 0; j < s->nprobes; j++) {
        v3 e = vsub(s->probe_pos[j], y);
        R d2 = vdot(e, e);
        if (d2 < best || (d2 == best && bi >= 0 && j < bi)) {
            best = d2;
            bi = j;
        }
    }
    return bi;
}

#define SSS_HITS 4
#define SSS_RADIUS_SIGMAS ((R)12.0f)

/* Keeps out[0..n) sorted by t with distinct t values, at most SSS_HITS: a
 * crossing at an existing t replaces it when its triangle index is lower. */
static void hits_insert(ohit* out, int* n, ohit c)
{
    for (int i = 0; i < *n; i++)
        if (out[i].t == c.t) {
            if (c.prim < out[i].prim) out[i] = c;
            return;
        }
    if (*n == SSS_HITS && !(c.t < out[*n - 1].t)) return;
    int i = *n < SSS_HITS ? (*n)++ : SSS_HITS - 1;
    while (i > 0 && out[i - 1].t > c.t) {
        out[i] = out[i - 1];
        i--;
    }
    out[i] = c;
}

/* Crossings of material m along o + t d, 0 < t < tmax (new: the importance-
 * sampled gather's probe ray): the SSS_HITS smallest distinct t, ascending, at
 * equal t the lower triangle index.  Every box overlapping the segment is
 * visited, so the result does not depend on traversal order. */
static int probe_collect(const oscene* s, v3 o, v3 d, R tmax, int m, ohit* out)
{
    int n = 0;
    if (s->nnodes == 0) return 0;
    double od[3] = {o.x, o.y, o.z};
    double inv[3] = {1.0 / (double)d.x, 1.0 / (double)d.y, 1.0 / (double)d.z};
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const bnode* b = &s->nodes[stack[--sp]];
        double tn;
        if (!box_hit(b, od, inv, 0.0, (double)tmax * (1.0 + 1e-6), &tn)) continue;
        if (b->left < 0) {
            for (int32_t k = 0; k < b->count; k++) {
                int32_t id = s->order[b->first + k];
                if (s->tri_mat[id] != m) continue;
                ohit c;
                if (!isect(o, d, s->p0[id], s->e1[id], s->e2[id], (R)0, tmax, &c.t, &c.b, &c.g)) continue;
                c.prim = id;
                hits_insert(out, &n, c);
            }
            continue;
        }
        stack[sp++] = b->right;
        stack[sp++] = b->left;
    }
    return n;
}

/* The 64-bit random state of the path whose TEA state `t` points to: every
 * path's state is an orc_rng and the shaders receive a pointer to its first
 * member (a pointer to a struct's first member converts back to the struct). */
static uint32_t* path_s64(uint32_t* t) { return ((orc_rng*)t)->s64; }

/* The coefficients that drive sampling for material P (the frozen ones of the
 * differentiated material, else P itself). */
static const osss* sss_sampling(const oscene* s, const osss* P)
{
    return s->sss_diff_mat >= 0 && P == &s->sss[s->sss_diff_mat] ? &s->sss_frozen : P;
}

/* Roulette probability of the extinction weight `beam` over distance t: its
 * channel mean, from the sampling coefficients. */
static R beam_prob(const oscene* s, const osss* P, R t, v3 beam)
{
    const osss* Q = sss_sampling(s, P);
    if (Q != P) {
        v3 ext = Q->dipole_model == 1 ? Q->deltaEddExtinction : Q->extinction;
        beam = V3(F_EXP((-t) * ext.x), F_EXP((-t) * ext.y), F_EXP((-t) * ext.z));
    }
    return (beam.x + beam.y + beam.z) / (R)3.0f;
}

/* The gather's draws come from the 64-bit stream, as the reference's gather
 * loop draws its roulette numbers (CU/subsurface_scattering_shader.cu:152-156). */
static v3 sss_gather(const oscene* s, const osss* P, v3 xo, v3 no, uint32_t* s64)
{
    v3 acc = V3(0, 0, 0);
    if (s->nprobes <= 0 || s->probe_samples <= 0) return acc;
    const osss* Q = sss_sampling(s, P); /* exit-point sampling and its pdf */
    v3 t1, t2;
    orc_create_onb(no, &t1, &t2);
    R smin = fmin(Q->transport.x, fmin(Q->transport.y, Q->transport.z));
    R r_max = SSS_RADIUS_SIGMAS / smin;
    for (int k = 0; k < s->probe_samples; k++) {
        R u = orc_rnd_accurate(s64), v = orc_rnd_accurate(s64), w = orc_rnd_accurate(s64), sx = orc_rnd_accurate(s64);
        int c = (int)((R)3.0f * u);
        if (c > 2) c = 2;
        R sig = c == 0 ? Q->transport.x : (c == 1 ? Q->transport.y : Q->transport.z);
        R r = -F_LOG((R)1.0f - v) / sig;
        if (!(r < r_max)) continue;
        R ang = (R)6.28318530717958647692f * w;
        R sp = F_SIN(ang), cp = F_COS(ang);
        R hh = sqrt(r_max * r_max - r * r);
        v3 top = vadd(vadd(xo, vscale(vadd(vscale(t1, cp), vscale(t2, sp)), r)), vscale(no, hh));
        v3 dir = vneg(no);
        /* crossings of the probes' material along the probe ray */
        ohit hits[SSS_HITS];
        int nh = probe_collect(s, top, dir, (R)2.0f * hh, s->probe_mat, hits);
        if (nh == 0) continue;
        int pick = (int)(sx * (R)nh);
        if (pick > nh - 1) pick = nh - 1;
        const ohit* hp = &hits[pick];
        v3 xi = vadd(top, vscale(dir, hp->t));
        v3 ni = hit_normal(s, hp);
        R cosn = fabs(vdot(ni, no));
        v3 dx = vsub(xi, xo);
        v3 inpl = vsub(dx, vscale(no, vdot(dx, no)));
        R rp = sqrt(vdot(inpl, inpl));
        R pr_r = (Q->transport.x * F_EXP(-Q->transport.x * rp) + Q->transport.y * F_EXP(-Q->transport.y * rp) +
                  Q->transport.z * F_EXP(-Q->transport.z * rp)) /
                 ((R)3.0f * (R)6.28318530717958647692f * rp);
        R pdf = pr_r * cosn / (R)nh;
        if (!(pdf > 0) || !isfinite(pdf)) continue;
        int j = nearest_probe(s, xi);
        if (j < 0) continue;
        v3 E = vscale(s->probe_L[j], s->probe_T12[j]);
        v3 S = P->dipole_model == 1 ? dirpole_bssrdf(xi, s->probe_sn[j], s->probe_w12[j], xo, no, P)
                                    : dipole_bssrdf(sqrt(vdot(dx, dx)), P);
        v3 es = vmul(E, S);
        acc = vadd(acc, V3(es.x / pdf, es.y / pdf, es.z / pdf));
    }
    return vmul(vscale(acc, (R)1.0f / (R)s->probe_samples), P->global_coeff);
}

static v3 shade_hit(const oscene* s, v3 o, v3 d, const ohit* h, int depth, int emit_light, uint32_t* seed,
                    R prev_pdf, ostats* st);

/* CU/subsurface_scattering_shader.cu:55-198 with the gather before the refracted path. */
static v3 shade_translucent(const oscene* s, int mi, v3 o, v3 d, const ohit* h, int depth, uint32_t* t, ostats* st)
{
    v3 zero = V3(0, 0, 0);
    if (depth > s->max_depth) return zero;
    const osss* P = &s->sss[mi];
    v3 xo = vadd(o, vscale(d, h->t));
    v3 no = hit_normal(s, h);
    v3 wo = vneg(d);
    R recip = (R)1.0f / P->relative_ior;
    uint32_t* t64 = path_s64(t); /* every draw of this program is 64-bit (RND_64, :8) */
    R reflect_xi = orc_rnd_accurate(t64);
    R co = vdot(wo, no);
    int inside = co < 0;
    v3 beam = V3(1, 1, 1);
    if (inside) {
        v3 ext = P->dipole_model == 1 ? P->deltaEddExtinction : P->extinction;
        beam = V3(F_EXP((-h->t) * ext.x), F_EXP((-h->t) * ext.y), F_EXP((-h->t) * ext.z));
        R prob = beam_prob(s, P, h->t, beam);
        if (orc_rnd_accurate(t64) >= prob) return zero;
        beam = V3(beam.x / prob, beam.y / prob, beam.z / prob);
        recip = P->relative_ior;
        no = vneg(no);
        co = -co;
    }
    R st2 = recip * recip * ((R)1.0f - co * co);
    R ct = 1, Rf = 1;
    if (st2 < (R)1.0f) {
        ct = sqrt((R)1.0f - st2);
        Rf = orc_fresnel_R3(co, ct, recip);
    }
    v3 result = zero;
    if (reflect_xi >= Rf) {
        if (!inside && mi == s->probe_mat) result = sss_gather(s, P, xo, no, t64);
        v3 wt = vsub(vscale(vsub(vscale(no, co), wo), recip), vscale(no, ct));
        result = vadd(result, radiance(s, xo, wt, depth + 1, 1, t, (R)-1, st));
    } else {
        v3 wr = vsub(vscale(no, (R)2.0f * co), wo);
        result = radiance(s, xo, wr, depth + 1, 1, t, (R)-1, st);
    }
    return vmul(result, beam);
}

/* The NEE reflection factor of the rough translucent boundary's Walter and
 * visible-normal models (CU/rough_subsurface_scattering_shader.cu:146-165):
 * returns G1(i, h) G1(l, h) D_wi(h) / (4 |i.n| |l.n|) with h the half vector of
 * i and l, and the Fresnel reflectance at l.h in *Rr. */
static R rough_nee_reflect(v3 w_i, v3 wl, v3 ffn, R ax, R ay, int dist, R recip, R* Rr)
{
    v3 rn = orc_normalize(vadd(w_i, wl));
    R cr = vdot(rn, wl);
    *Rr = 1;
    R sr2 = recip * recip * ((R)1.0f - cr * cr);
    if (sr2 < (R)1.0f) *Rr = orc_fresnel_R3(cr, sqrt((R)1.0f - sr2), recip);
    R gi = masking_xy(w_i, rn, ffn, ax, ay, dist);
    R go = masking_xy(wl, rn, ffn, ax, ay, dist);
    R D = orc_eval_visible_normal(w_i, rn, ffn, ax, ay, dist);
    return gi * go * D / ((R)4.0f * fabs(vdot(w_i, ffn) * vdot(wl, ffn)));
}

/* Batch of rough_nee_reflect in the frame n = +z: 3 floats per direction;
 * out = Rr * factor per pair. */
void orc_rough_nee_reflect_batch(int n, const float* wi, const float* wl, float ax, float ay, int dist, float recip,
                                 double* out)
{
    for (int k = 0; k < n; k++) {
        R Rr;
        R f = rough_nee_reflect(V3(wi[3 * k], wi[3 * k + 1], wi[3 * k + 2]), V3(wl[3 * k], wl[3 * k + 1], wl[3 * k + 2]),
                                V3(0, 0, 1), ax, ay, dist, recip, &Rr);
        out[k] = (double)(Rr * f);
    }
}

/* CU/rough_subsurface_scattering_shader.cu:64-455: microfacet boundary (Walter,
 * visible-normal or multiple-scattering model) over the diffusion term.  The
 * NEE BSDF factor is evaluated before the shadow test (the product traces that
 * segment after shading), so a multiple-scattering estimate draws its random
 * numbers whether or not the light is visible. */
static v3 shade_rough_translucent(const oscene* s, int mi, v3 o, v3 d, const ohit* h, int depth, int emit_light,
                                  uint32_t* t, ostats* st)
{
    v3 zero = V3(0, 0, 0);
    if (depth > s->max_depth) return zero;
    const osss* P = &s->sss[mi];
    const orc_material* M = &s->mats[mi];
    v3 p = vadd(o, vscale(d, h->t));
    v3 normal = hit_normal(s, h);
    v3 ffn = orc_faceforward(normal, vneg(d), normal);
    v3 w_i = vneg(d);
    R ax = M->roughness[0], ay = M->roughness[1];
    int dist = M->normal_distribution, model = M->microfacet_model;
    R recip = (R)1.0f / P->relative_ior;
    /* 64-bit draws (RND_64): reflect_xi, z1, z2 (:81-84), the light pick
     * (:108-109), the extinction roulette (:186-187) and the gather; TEA: the
     * light sample and the multiple-scattering walk and evaluation */
    uint32_t* t64 = path_s64(t);
    R reflect_xi = orc_rnd_accurate(t64);
    R z1 = orc_rnd_accurate(t64), z2 = orc_rnd_accurate(t64);
    v3 mn;
    if (model == 0) {
        mn = orc_sample_normal_walter(ffn, ax, z1, z2, dist); /* the reference's argument order (:96) */
    } else {
        v3 l = orc_to_local(w_i, ffn);
        mn = orc_to_world(dist == 1 ? orc_ggx_sample_vndf(l, ax, ay, z1, z2) : orc_beckmann_sample_vndf(l, ax, ay, z1, z2),
                          ffn);
    }
    R cos_i = vdot(w_i, mn);
    int inside = vdot(w_i, normal) < 0;
    /* NEE over the reference's lights (the environment entry is not one of them) */
    int nreal = s->nlights - (s->env_is ? 1 : 0);
    uint32_t li = (uint32_t)((R)nreal * orc_rnd_accurate(t64));
    v3 result = zero;
    int refl_emit = emit_light;
    if (nreal > 0) {
        if (li >= (uint32_t)nreal) li = (uint32_t)nreal - 1;
        R light_pdf = (R)1 / (R)nreal;
        v3 wl, rad;
        R ldist, penv;
        direct_light(s, &s->lights[li], p, t, &wl, &rad, &ldist, &penv);
        rad = V3(rad.x / light_pdf, rad.y / light_pdf, rad.z / light_pdf);
        R cl = vdot(ffn, wl);
        if (!inside && cl > 0) {
            v3 c;
            if (model == 2) {
                v3 lwi = orc_to_local(w_i, ffn), lwl = orc_to_local(wl, ffn);
                R f = dist == 1 ? orc_ggx_ms_dielectric_eval(lwi, lwl, recip, ax, ay, t, MS_WALK_CAP)
                                : orc_beckmann_ms_dielectric_eval(lwi, lwl, recip, ax, ay, t, MS_WALK_CAP);
                c = vscale(rad, f);
            } else {
                R Rr;
                R wr = rough_nee_reflect(w_i, wl, ffn, ax, ay, dist, recip, &Rr);
                c = vscale(vscale(rad, Rr), wr);
            }
            ohit sh;
            st->shd++;
            if (!trace(s, p, wl, s->eps, ldist, 1, &sh)) {
                result = c;
                if (model != 2) refl_emit = 1; /* :165 */
            }
        }
    }
    v3 beam = V3(1, 1, 1);
    v3 nd;
    R weight = 1;
    int refract;
    if (model == 2) {
        v3 lo;
        int order;
        v3 lwi = orc_to_local(w_i, normal);
        R eta = (R)1.0f / P->relative_ior;
        if (!(dist == 1 ? orc_ggx_ms_dielectric_sample(lwi, eta, ax, ay, t, MS_WALK_CAP, &lo, &order)
                        : orc_beckmann_ms_dielectric_sample(lwi, eta, ax, ay, t, MS_WALK_CAP, &lo, &order)))
            return result;
        nd = orc_to_world(lo, normal);
        refract = vdot(nd, ffn) < 0;
    }
    if (inside) {
        v3 ext = P->dipole_model == 1 ? P->deltaEddExtinction : P->extinction;
        beam = V3(F_EXP((-h->t) * ext.x), F_EXP((-h->t) * ext.y), F_EXP((-h->t) * ext.z));
        R prob = beam_prob(s, P, h->t, beam);
        if (orc_rnd_accurate(t64) >= prob) return result;
        beam = V3(beam.x / prob, beam.y / prob, beam.z / prob);
        recip = P->relative_ior;
    }
    v3 no = inside ? vneg(normal) : normal;
    if (model != 2) {
        R st2 = recip * recip * ((R)1.0f - cos_i * cos_i);
        R cos_t = 1, Rf = 1;
        if (st2 < (R)1.0f) {
            cos_t = sqrt((R)1.0f - st2);
            Rf = orc_fresnel_R3(cos_i, cos_t, recip);
        }
        refract = reflect_xi >= Rf;
        nd = refract ? vsub(vscale(vsub(vscale(mn, cos_i), w_i), recip), vscale(mn, cos_t))
                     : vsub(vscale(mn, (R)2.0f * cos_i), w_i);
        if (model == 0) {
            R gi = orc_masking_iso(w_i, mn, ffn, ax, dist);
            R go = orc_masking_iso(nd, mn, ffn, ax, dist);
            weight = fabs(vdot(w_i, mn)) * gi * go / (fabs(vdot(w_i, ffn)) * fabs(vdot(ffn, mn)));
        } else {
            weight = masking_xy(nd, mn, ffn, ax, ay, dist);
        }
    }
    v3 acc = zero;
    int child_emit = 1;
    if (refract) {
        if (!inside && mi == s->probe_mat) acc = vscale(sss_gather(s, P, p, no, t64), weight);
    } else if (model != 2) {
        child_emit = refl_emit;
    }
    v3 child = radiance(s, p, nd, depth + 1, child_emit, t, (R)-1, st);
    acc = vadd(acc, vscale(child, weight));
    return vadd(result, vmul(acc, beam));
}

static v3 radiance(const oscene* s, v3 o, v3 d, int depth, int emit_light, uint32_t* seed, R prev_pdf, ostats* st)
{
    ohit h;
    st->ext++;
    if (!trace(s, o, d, s->eps, RAY_MAX, 0, &h)) {
        if (!s->has_env) return V3(0, 0, 0);
        v3 E = env_lookup(s, d);
        if (prev_pdf > 0 && s->env_is) E = vscale(E, mis_power(prev_pdf, env_pdf(s, d) / (R)s->nlights));
        return E;
    }
    return shade_hit(s, o, d, &h, depth, emit_light, seed, prev_pdf, st);
}

static v3 shade_hit(const oscene* s, v3 o, v3 d, const ohit* hp, int depth, int emit_light, uint32_t* seed,
                    R prev_pdf, ostats* st)
{
    ohit h = *hp;
    (void)prev_pdf;
    const orc_material* m = &s->mats[s->tri_mat[h.prim]];
    if (m->kind == ORC_EMITTER) {
        /* CU/arealight_shader.cu:36-50 */
        v3 normal = hit_normal(s, &h);
        if (vdot(normal, vneg(d)) > 0 && emit_light) return V3(m->radiance[0], m->radiance[1], m->radiance[2]);
        return V3(0, 0, 0);
    }
    if (m->kind == ORC_LAMBERT) return shade_lambert(s, m, o, d, &h, depth, &emit_light, seed, st);
    if (m->kind == ORC_CONDUCTOR) return shade_conductor(s, m, &s->eta[s->tri_mat[h.prim]], o, d, &h, depth, seed, st);
    if (m->kind == ORC_DIELECTRIC) return shade_dielectric(s, m, o, d, &h, depth, seed, st);
    if (m->kind == ORC_TRANSPARENT) return shade_transparent(s, m, o, d, &h, depth, seed, st);
    if (m->kind == ORC_FLAT) return shade_flat(s, m, o, d, &h, depth, seed, st);
    if (m->kind == ORC_ROUGH_DIFFUSE) return shade_rough_diffuse(s, m, o, d, &h, depth, emit_light, seed, st);
    if (m->kind == ORC_NORMAL) {
        /* CU/normal_shader.cu:25-29: normal * 0.5 + 0.5 */
        v3 n = hit_normal(s, &h);
        return V3(n.x * (R)0.5f + (R)0.5f, n.y * (R)0.5f + (R)0.5f, n.z * (R)0.5f + (R)0.5f);
    }
    if (m->kind == ORC_INTERFACE) return shade_interface(s, m, o, d, &h, depth, emit_light, seed, st);
    if (m->kind == ORC_TRANSLUCENT) return shade_translucent(s, s->tri_mat[h.prim], o, d, &h, depth, seed, st);
    if (m->kind == ORC_ROUGH_TRANSLUCENT)
        return shade_rough_translucent(s, s->tri_mat[h.prim], o, d, &h, depth, emit_light, seed, st);
    return V3(0, 0, 0);
}

/* path_tracer (CU/path_tracer.cu:32-74) for one (pixel, frame). */
static v3 pixel_sample_cam(const oscene* s, const ocam* c, int W, int H, int x, int y, uint32_t frame, ostats* st)
{
    /* prd.seed and prd.seed64 (CU/path_tracer.cu:48-49) */
    orc_rng r;
    r.tea = orc_tea16((uint32_t)W * (uint32_t)y + (uint32_t)x, frame);
    r.s64[0] = r.s64[1] = r.tea;
    R jx = orc_rnd_tea(&r.tea);
    R jy = orc_rnd_tea(&r.tea);
    R ipx = ((R)x + jx) / (R)W * (R)2.0f - (R)1.0f;
    R ipy = ((R)y + jy) / (R)H * (R)2.0f - (R)1.0f;
    v3 dir = orc_normalize(vadd(vadd(vscale(c->U, ipx), vscale(c->V, ipy)), c->W));
    return radiance(s, c->eye, dir, 0, 1, &r.tea, (R)-1, st);
}

static v3 pixel_sample(const oscene* s, int W, int H, int x, int y, uint32_t frame, ostats* st)
{
    ocam c = {s->eye, s->U, s->V, s->W};
    return pixel_sample_cam(s, &c, W, H, x, y, frame, st);
}

/* ---------------------------------------------------------------- API */

void* orc_scene_create(const orc_scene_desc* d)
{
    oscene* s = (oscene*)calloc(1, sizeof(oscene));
    s->sss_diff_mat = -1;
    s->ntris = d->ntris;
    s->p0 = (v3*)malloc(sizeof(v3) * (size_t)(d->ntris + 1));
    s->e1 = (v3*)malloc(sizeof(v3) * (size_t)(d->ntris + 1));
    s->e2 = (v3*)malloc(sizeof(v3) * (size_t)(d->ntris + 1));
    for (int64_t t = 0; t < d->ntris; t++) {
        const int32_t* ti = d->idx + 3 * t;
        /* e1 = p1 - p0, e2 = p2 - p0 rounded in float, as the product stores them */
        float p[3][3];
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) p[k][a] = d->verts[3 * ti[k] + a];
        s->p0[t] = V3(p[0][0], p[0][1], p[0][2]);
        s->e1[t] = V3(p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]);
        s->e2[t] = V3(p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]);
    }
    if (d->nidx && d->normals) {
        s->nidx = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(d->ntris + 1));
        memcpy(s->nidx, d->nidx, sizeof(int32_t) * 3 * (size_t)d->ntris);
        s->normals = (v3*)malloc(sizeof(v3) * (size_t)(d->nnormals + 1));
        for (int64_t k = 0; k < d->nnormals; k++)
            s->normals[k] = V3(d->normals[3 * k], d->normals[3 * k + 1], d->normals[3 * k + 2]);
    }
    /* disk and sphere emitters, in light order; their materials follow the
     * triangles' in tri_mat */
    int32_t nan = 0;
    for (int i = 0; i < d->nlights; i++) nan += d->lights[i].type == ORC_DISK || d->lights[i].type == ORC_SPHERE;
    s->nan = nan;
    s->an_c = (v3*)malloc(sizeof(v3) * (size_t)(nan + 1));
    s->an_n = (v3*)malloc(sizeof(v3) * (size_t)(nan + 1));
    s->an_r = (R*)malloc(sizeof(R) * (size_t)(nan + 1));
    s->an_sphere = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nan + 1));
    s->tri_mat = (int32_t*)malloc(sizeof(int32_t) * (size_t)(d->ntris + nan + 1));
    memcpy(s->tri_mat, d->tri_mat, sizeof(int32_t) * (size_t)d->ntris);
    for (int i = 0, k = 0; i < d->nlights; i++) {
        const orc_light* l = &d->lights[i];
        if (l->type != ORC_DISK && l->type != ORC_SPHERE) continue;
        s->an_c[k] = V3(l->params[0], l->params[1], l->params[2]);
        s->an_r[k] = l->shape[0];
        s->an_sphere[k] = l->type == ORC_SPHERE;
        /* DiskAreaLight::computeNormal (ORF/Light.cpp:469-474) */
        R th = (R)l->shape[1] * PI_R / (R)180.0f, ph = (R)l->shape[2] * PI_R / (R)180.0f;
        R sth = F_SIN(th);
        s->an_n[k] = V3(sth * F_SIN(ph), F_COS(th), sth * F_COS(ph));
        s->tri_mat[d->ntris + k] = l->material;
        k++;
    }
    s->nmats = d->nmats;
    s->mats = (orc_material*)malloc(sizeof(orc_material) * (size_t)(d->nmats + 1));
    memcpy(s->mats, d->mats, sizeof(orc_material) * (size_t)d->nmats);
    s->eta = (cplx3*)malloc(sizeof(cplx3) * (size_t)(d->nmats + 1));
    for (int i = 0; i < d->nmats; i++) {
        /* eta = n1 / ior (CU/metallic_shader.cu:59-61) */
        cplx c0 = {d->mats[i].ior_re[0], d->mats[i].ior_im[0]}, c1 = {d->mats[i].ior_re[1], d->mats[i].ior_im[1]},
             c2 = {d->mats[i].ior_re[2], d->mats[i].ior_im[2]};
        s->eta[i].x = orc_cdiv_fc((R)1.0f, c0);
        s->eta[i].y = orc_cdiv_fc((R)1.0f, c1);
        s->eta[i].z = orc_cdiv_fc((R)1.0f, c2);
    }
    s->ieta = (cplx3*)malloc(sizeof(cplx3) * (size_t)(d->nmats + 1));
    for (int i = 0; i < d->nmats; i++) {
        s->ieta[i].x = (cplx){d->mats[i].ior_re[0], d->mats[i].ior_im[0]};
        s->ieta[i].y = (cplx){d->mats[i].ior_re[1], d->mats[i].ior_im[1]};
        s->ieta[i].z = (cplx){d->mats[i].ior_re[2], d->mats[i].ior_im[2]};
    }
    s->sss = (osss*)calloc((size_t)(d->nmats + 1), sizeof(osss));
    for (int i = 0; i < d->nmats; i++) s->sss[i] = sss_coefficients(&d->mats[i]);
    s->probe_mat = -1;
    s->has_env = d->env != NULL && d->env_w > 0 && d->env_h > 0;
    s->env_is = s->has_env && d->env_is;
    int extra = s->env_is ? 1 : 0;
    s->nlights = d->nlights + extra;
    s->lights = (orc_light*)calloc((size_t)(s->nlights + 1), sizeof(orc_light));
    memcpy(s->lights, d->lights, sizeof(orc_light) * (size_t)d->nlights);
    if (extra) s->lights[d->nlights].type = ORC_ENV;
    s->trilights = (orc_trilight*)malloc(sizeof(orc_trilight) * (size_t)(d->ntrilights + 1));
    memcpy(s->trilights, d->trilights, sizeof(orc_trilight) * (size_t)d->ntrilights);
    if (s->has_env) {
        s->env_w = d->env_w;
        s->env_h = d->env_h;
        s->env = (float*)malloc(sizeof(float) * 4 * (size_t)d->env_w * d->env_h);
        memcpy(s->env, d->env, sizeof(float) * 4 * (size_t)d->env_w * d->env_h);
        if (s->env_is) build_env_tables(s);
    }
    s->eye = V3(d->eye[0], d->eye[1], d->eye[2]);
    s->U = V3(d->U[0], d->U[1], d->U[2]);
    s->V = V3(d->V[0], d->V[1], d->V[2]);
    s->W = V3(d->W[0], d->W[1], d->W[2]);
    s->max_depth = d->max_depth;
    s->eps = d->scene_epsilon;
    build_bvh(s);
    return s;
}

void orc_scene_destroy(void* p)
{
    oscene* s = (oscene*)p;
    if (!s) return;
    free(s->p0);
    free(s->e1);
    free(s->e2);
    free(s->nidx);
    free(s->normals);
    free(s->tri_mat);
    free(s->ieta);
    free(s->an_c);
    free(s->an_n);
    free(s->an_r);
    free(s->an_sphere);
    free(s->nodes);
    free(s->order);
    free(s->mats);
    free(s->eta);
    free(s->lights);
    free(s->trilights);
    free(s->env);
    free(s->mpdf);
    free(s->mcdf);
    free(s->cpdf);
    free(s->ccdf);
    free(s->alb_eval);
    free(s->alb_frozen);
    free(s->rough_eval);
    free(s->rough_frozen);
    free(s->sss);
    free(s->probe_pos);
    free(s->probe_nrm);
    free(s->probe_L);
    free(s->probe_w12);
    free(s->probe_T12);
    free(s->probe_sn);
    free(s);
}

void orc_render(void* p, int W, int H, int x0, int y0, int x1, int y1, uint32_t frame0, uint32_t spp, double* out,
                uint64_t* counters, int threads)
{
    const oscene* s = (const oscene*)p;
    int rw = x1 - x0, rh = y1 - y0;
    uint64_t bad = 0, ext = 0, shd = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads) reduction(+ : bad, ext, shd)
    for (int64_t i = 0; i < (int64_t)rw * rh; i++) {
        int x = x0 + (int)(i % rw), y = y0 + (int)(i / rw);
        double acc[4] = {0, 0, 0, 0};
        ostats st = {0, 0};
        for (uint32_t f = frame0; f < frame0 + spp; f++) {
            v3 L = pixel_sample(s, W, H, x, y, f, &st);
            if (isfinite(L.x) && isfinite(L.y) && isfinite(L.z)) {
                acc[0] += (double)L.x;
                acc[1] += (double)L.y;
                acc[2] += (double)L.z;
                acc[3] += 1.0;
            } else {
                bad++;
            }
        }
        ext += st.ext;
        shd += st.shd;
        memcpy(out + 4 * i, acc, sizeof(acc));
    }
    if (counters) {
        counters[0] = bad;
        counters[1] = ext;
        counters[2] = shd;
        counters[3] = (uint64_t)rw * rh * spp;
    }
}

/* DepthTracer (CU/depth_tracer.cu:33-58 + CU/depth_shader.cu:13-19): per pixel
 * of `frame`, (1 + n) / 2 of the face-forwarded shading normal and the hit
 * distance; 0.5 grey and 0 where the ray leaves the scene or meets a disk or
 * sphere emitter (no depth program on their materials).  out: 4 per pixel. */
void orc_depth_samples(void* p, int W, int H, const int32_t* px, const int32_t* py, int n, uint32_t frame,
                       double* out)
{
    const oscene* s = (const oscene*)p;
    for (int i = 0; i < n; i++) {
        int x = px[i], y = py[i];
        uint32_t seed = orc_tea16((uint32_t)W * (uint32_t)y + (uint32_t)x, frame);
        R jx = orc_rnd_tea(&seed);
        R jy = orc_rnd_tea(&seed);
        R ipx = ((R)x + jx) / (R)W * (R)2.0f - (R)1.0f;
        R ipy = ((R)y + jy) / (R)H * (R)2.0f - (R)1.0f;
        v3 dir = orc_normalize(vadd(vadd(vscale(s->U, ipx), vscale(s->V, ipy)), s->W));
        ohit h;
        v3 nrm = V3(0, 0, 0);
        R t = 0;
        if (trace(s, s->eye, dir, s->eps, RAY_MAX, 0, &h) && h.prim < s->ntris) {
            v3 nn = hit_normal(s, &h);
            nrm = orc_faceforward(nn, vneg(dir), nn);
            t = h.t;
        }
        out[4 * i] = (double)(((R)1.0f + nrm.x) * (R)0.5f);
        out[4 * i + 1] = (double)(((R)1.0f + nrm.y) * (R)0.5f);
        out[4 * i + 2] = (double)(((R)1.0f + nrm.z) * (R)0.5f);
        out[4 * i + 3] = (double)t;
    }
}

void orc_render_samples(void* p, int W, int H, const int32_t* px, const int32_t* py, int n, uint32_t frame,
                        double* out, int threads)
{
    const oscene* s = (const oscene*)p;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int i = 0; i < n; i++) {
        ostats st = {0, 0};
        v3 L = pixel_sample(s, W, H, px[i], py[i], frame, &st);
        out[3 * i] = (double)L.x;
        out[3 * i + 1] = (double)L.y;
        out[3 * i + 2] = (double)L.z;
    }
}

void orc_env_tables(void* p, double* mpdf, double* mcdf, double* cpdf, double* ccdf)
{
    const oscene* s = (const oscene*)p;
    if (!s->env_is) return;
    size_t w = s->env_w, h = s->env_h;
    for (size_t j = 0; j < h; j++) {
        mpdf[j] = (double)s->mpdf[j];
        mcdf[j] = (double)s->mcdf[j];
    }
    for (size_t k = 0; k < w * h; k++) {
        cpdf[k] = (double)s->cpdf[k];
        ccdf[k] = (double)s->ccdf[k];
    }
}

void orc_closest_hits(void* p, int n, const float* o, const float* d, float tmin, float* t, int32_t* prim,
                      int brute_force)
{
    const oscene* s = (const oscene*)p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
        ohit h;
        v3 oo = V3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), dd = V3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        if (brute_force) trace_brute(s, oo, dd, (R)tmin, RAY_MAX, &h);
        else trace(s, oo, dd, (R)tmin, RAY_MAX, 0, &h);
        t[i] = (float)h.t;
        prim[i] = h.prim;
    }
}

/* Draws of env_sample (one TEA stream per draw, seeded tea16(i, seed)) with
 * their returned pdf and the pdf env_pdf evaluates at the drawn direction. */
void orc_env_sample_batch(void* p, uint32_t seed, int n, double* dir, double* pdf_sample, double* pdf_eval)
{
    const oscene* s = (const oscene*)p;
    if (!s->env_is) return;
    for (int i = 0; i < n; i++) {
        uint32_t t = orc_tea16((uint32_t)i, seed);
        v3 d, L;
        R pdf;
        env_sample(s, &t, &d, &L, &pdf);
        dir[3 * i] = (double)d.x;
        dir[3 * i + 1] = (double)d.y;
        dir[3 * i + 2] = (double)d.z;
        pdf_sample[i] = (double)pdf;
        pdf_eval[i] = (double)env_pdf(s, d);
    }
}

/* ---------------------------------------------------------------- differentiable form (new) */

static R* dup_tex(const double* src, size_t n)
{
    if (!src) return NULL;
    R* d = (R*)malloc(sizeof(R) * (n + 1));
    for (size_t i = 0; i < n; i++) d[i] = (R)src[i];
    return d;
}

/* Textures of the differentiable form.  albedo_* are tex_h x tex_w x 3, rough_*
 * tex_h x tex_w; *_frozen (NULL: same as *_eval) drive sampling and Russian
 * roulette, *_eval the BSDF weights and NEE, so the estimator is a smooth
 * function of *_eval at fixed random numbers (what a finite difference of the
 * detached-sampling adjoint differentiates).  mat < 0 disables a texture. */
void orc_diff_set(void* p, int tex_w, int tex_h, int alb_mat, int alb_uv, const double* alb_uvp, int rough_mat,
                  int rough_uv, const double* rough_uvp, const double* albedo_eval, const double* albedo_frozen,
                  const double* rough_eval, const double* rough_frozen)
{
    oscene* s = (oscene*)p;
    free(s->alb_eval);
    free(s->alb_frozen);
    free(s->rough_eval);
    free(s->rough_frozen);
    size_t n = (size_t)tex_w * (size_t)tex_h;
    s->tex_w = tex_w;
    s->tex_h = tex_h;
    s->alb_mat = alb_mat;
    s->rough_mat = rough_mat;
    s->alb_uv = alb_uv;
    s->rough_uv = rough_uv;
    for (int k = 0; k < 4; k++) {
        s->alb_uvp[k] = (R)alb_uvp[k];
        s->rough_uvp[k] = (R)rough_uvp[k];
    }
    s->alb_eval = alb_mat >= 0 ? dup_tex(albedo_eval, 3 * n) : NULL;
    s->alb_frozen = alb_mat >= 0 ? dup_tex(albedo_frozen ? albedo_frozen : albedo_eval, 3 * n) : NULL;
    s->rough_eval = rough_mat >= 0 ? dup_tex(rough_eval, n) : NULL;
    s->rough_frozen = rough_mat >= 0 ? dup_tex(rough_frozen ? rough_frozen : rough_eval, n) : NULL;
    s->diff_on = (s->alb_eval || s->rough_eval) && n > 0;
}

/* Views: cams holds 12 floats per view (eye, U, V, W).  images (may be NULL)
 * receives per-view mean images over the finite samples of frames
 * frame0[v] + shift + s, s < spp; the return value is
 * sum over views, pixels, channels of dimg * (sum of finite samples) / spp,
 * the linear loss whose gradient the adjoint accumulates. */
double orc_diff_eval(void* p, int W, int H, int nviews, const float* cams, const uint32_t* frame0, uint32_t spp,
                     uint32_t shift, const double* dimg, double* images, int threads)
{
    const oscene* s = (const oscene*)p;
    double loss = 0;
    int64_t npx = (int64_t)W * H;
    for (int v = 0; v < nviews; v++) {
        const float* c = cams + 12 * v;
        ocam cam = {V3(c[0], c[1], c[2]), V3(c[3], c[4], c[5]), V3(c[6], c[7], c[8]), V3(c[9], c[10], c[11])};
        double part = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads) reduction(+ : part)
        for (int64_t i = 0; i < npx; i++) {
            int x = (int)(i % W), y = (int)(i / W);
            double acc[3] = {0, 0, 0}, cnt = 0;
            ostats st = {0, 0};
            for (uint32_t f = 0; f < spp; f++) {
                v3 L = pixel_sample_cam(s, &cam, W, H, x, y, frame0[v] + shift + f, &st);
                if (isfinite(L.x) && isfinite(L.y) && isfinite(L.z)) {
                    acc[0] += (double)L.x;
                    acc[1] += (double)L.y;
                    acc[2] += (double)L.z;
                    cnt += 1;
                }
            }
            size_t o = 3 * ((size_t)v * npx + i);
            if (dimg) part += (dimg[o] * acc[0] + dimg[o + 1] * acc[1] + dimg[o + 2] * acc[2]) / (double)spp;
            if (images)
                for (int k = 0; k < 3; k++) images[o + k] = cnt > 0 ? acc[k] / cnt : 0.0;
        }
        loss += part;
    }
    return loss;
}

/* Differentiable scattering coefficients of material m (new): eval6 and frozen6
 * are (absorption RGB, scattering RGB) in 1/mm; the material shades with eval6
 * and samples with frozen6 (NULL: eval6).  m < 0 restores every material's own
 * coefficients.  The probes keep the values they hold. */
void orc_diff_set_sss(void* p, int m, const double* eval6, const double* frozen6)
{
    oscene* s = (oscene*)p;
    if (s->sss_diff_mat >= 0) s->sss[s->sss_diff_mat] = sss_coefficients(&s->mats[s->sss_diff_mat]);
    s->sss_diff_mat = -1;
    if (m < 0 || !eval6) return;
    const double* f = frozen6 ? frozen6 : eval6;
    s->sss[m] = sss_coefficients_with(&s->mats[m], eval6, eval6 + 3);
    s->sss_frozen = sss_coefficients_with(&s->mats[m], f, f + 3);
    s->sss_diff_mat = m;
}

/* ---------------------------------------------------------------- translucent API */

void orc_probes_set(void* p, int n, const float* pos, const float* nrm, float cell, int samples, int material)
{
    oscene* s = (oscene*)p;
    free(s->probe_pos);
    free(s->probe_nrm);
    free(s->probe_L);
    free(s->probe_w12);
    free(s->probe_T12);
    free(s->probe_sn);
    s->nprobes = n;
    s->probe_cell = (R)cell;
    s->probe_samples = samples;
    s->probe_mat = material;
    s->probe_pos = (v3*)calloc((size_t)n + 1, sizeof(v3));
    s->probe_nrm = (v3*)calloc((size_t)n + 1, sizeof(v3));
    s->probe_L = (v3*)calloc((size_t)n + 1, sizeof(v3));
    s->probe_w12 = (v3*)calloc((size_t)n + 1, sizeof(v3));
    s->probe_T12 = (R*)calloc((size_t)n + 1, sizeof(R));
    s->probe_sn = (v3*)calloc((size_t)n + 1, sizeof(v3));
    for (int j = 0; j < n; j++) {
        s->probe_pos[j] = V3(pos[4 * j], pos[4 * j + 1], pos[4 * j + 2]);
        s->probe_nrm[j] = V3(nrm[4 * j], nrm[4 * j + 1], nrm[4 * j + 2]);
        s->probe_sn[j] = s->probe_nrm[j];
    }
}

void orc_probes_values(void* p, const float* L, const float* w12, const float* sn)
{
    oscene* s = (oscene*)p;
    for (int j = 0; j < s->nprobes; j++) {
        s->probe_L[j] = V3(L[4 * j], L[4 * j + 1], L[4 * j + 2]);
        s->probe_w12[j] = V3(w12[4 * j], w12[4 * j + 1], w12[4 * j + 2]);
        s->probe_T12[j] = (R)w12[4 * j + 3];
        if (sn) s->probe_sn[j] = V3(sn[4 * j], sn[4 * j + 1], sn[4 * j + 2]);
    }
}

/* Light that has bounced at least once before reaching `o` along `d`: the first
 * segment's emitter and (when the environment is a sampled light) environment
 * belong to the light-source strategy. */
static v3 probe_indirect(const oscene* s, v3 o, v3 d, uint32_t* t, ostats* st)
{
    ohit h;
    st->ext++;
    if (!trace(s, o, d, s->eps, RAY_MAX, 0, &h)) {
        if (!s->has_env || s->env_is) return V3(0, 0, 0);
        return env_lookup(s, d);
    }
    return shade_hit(s, o, d, &h, 0, 0, t, (R)-1, st);
}

/* Microfacet normal of a rough boundary for a probe lit from wi
 * (CU/sample_camera.cu:130-151, :173-191; the Walter branch keeps the
 * reference's (a_x, z1, z2) argument order). */
static v3 probe_facet(const orc_material* M, v3 wi, v3 n, uint32_t* t64)
{
    R z1 = orc_rnd_accurate(t64);
    R z2 = orc_rnd_accurate(t64);
    R ax = M->roughness[0], ay = M->roughness[1];
    int dist = M->normal_distribution;
    if (M->microfacet_model == 0) return orc_sample_normal_walter(n, ax, z1, z2, dist);
    v3 l = orc_to_local(wi, n);
    v3 m = dist == 1 ? orc_ggx_sample_vndf(l, ax, ay, z1, z2) : orc_beckmann_sample_vndf(l, ax, ay, z1, z2);
    return orc_to_world(m, n);
}

/* CU/sample_camera.cu:109-268 for one probe. */
static void probe_sample(const oscene* s, const osss* P, const orc_material* M, int j, uint32_t frame, v3* L_out,
                         v3* w12_out, R* T12_out, v3* sn_out)
{
    v3 pos = s->probe_pos[j], n = s->probe_nrm[j];
    int rough = M->kind == ORC_ROUGH_TRANSLUCENT;
    /* t and t64 as CU/sample_camera.cu:54-57 seeds them; the 64-bit draws
     * that place the reference's probe (:58, :77-78) are drawn and unused */
    orc_rng r;
    r.tea = orc_tea16((uint32_t)j, frame);
    r.s64[0] = r.s64[1] = r.tea;
    uint32_t* t = &r.tea;
    uint32_t* t64 = r.s64;
    for (int k = 0; k < 3; k++) orc_rnd_accurate(t64);
    R recip = (R)1.0f / P->relative_ior;
    v3 Le = V3(0, 0, 0), wi = n, sn = n;
    ostats st = {0, 0};
    if (orc_rnd(t) < (R)0.5f) { /* rnd(t), the LCG draw (:116) */
        uint32_t li = (uint32_t)((R)s->nlights * orc_rnd_accurate(t64));
        if (s->nlights > 0) {
            if (li >= (uint32_t)s->nlights) li = (uint32_t)s->nlights - 1;
            R light_pdf = (R)1 / (R)s->nlights;
            v3 wl, rad;
            R dist, penv;
            direct_light(s, &s->lights[li], pos, t, &wl, &rad, &dist, &penv);
            wi = wl;
            if (rough) sn = probe_facet(M, wi, n, t64);
            R cs = fmax(vdot(wl, n), (R)0);
            if (cs > 0) {
                ohit sh;
                if (!trace(s, pos, wl, s->eps, dist, 1, &sh)) {
                    v3 a = vscale(rad, cs);
                    Le = V3(a.x / light_pdf, a.y / light_pdf, a.z / light_pdf);
                }
            }
        } else if (rough) {
            sn = probe_facet(M, wi, n, t64);
        }
    } else {
        wi = orc_sample_cosine_weighted(n, t);
        if (rough) sn = probe_facet(M, wi, n, t64);
        Le = vscale(probe_indirect(s, pos, wi, t, &st), PI_R);
    }
    Le = V3(Le.x / (R)0.5f, Le.y / (R)0.5f, Le.z / (R)0.5f);
    R ci = fmax(vdot(wi, sn), (R)0);
    R st2 = recip * recip * ((R)1.0f - ci * ci);
    R ct = sqrt((R)1.0f - st2);
    v3 w12 = vsub(vscale(vsub(vscale(sn, ci), wi), recip), vscale(sn, ct));
    R T12 = (R)1.0f - orc_fresnel_R3(ci, ct, recip);
    R weight = T12;
    if (rough && M->microfacet_model != 2) {
        /* CU/sample_camera.cu:240-254: the boundary's transmission weight */
        R ax = M->roughness[0], ay = M->roughness[1];
        int dist = M->normal_distribution;
        R go = masking_xy(w12, sn, n, ax, ay, dist);
        R w = go;
        if (M->microfacet_model == 0) {
            R gi = masking_xy(wi, sn, n, ax, ay, dist);
            w = fabs(vdot(wi, sn)) * gi / (fabs(vdot(wi, n)) * fabs(vdot(n, sn))) * go;
        }
        weight = w * T12;
    }
    *w12_out = w12;
    *T12_out = weight;
    *sn_out = sn;
    *L_out = Le;
}

void orc_probe_pass(void* p, uint32_t frame, float* L, float* w12, float* sn, int threads)
{
    oscene* s = (oscene*)p;
    if (s->nprobes <= 0 || s->probe_mat < 0) return;
    oscene off = *s;
    off.probe_mat = -1; /* no gathers while the probes are being written */
    const osss* P = &s->sss[s->probe_mat];
    const orc_material* M = &s->mats[s->probe_mat];
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (int j = 0; j < s->nprobes; j++) {
        v3 Lj, wj, nj;
        R Tj;
        probe_sample(&off, P, M, j, frame, &Lj, &wj, &Tj, &nj);
        s->probe_L[j] = Lj;
        s->probe_w12[j] = wj;
        s->probe_T12[j] = Tj;
        s->probe_sn[j] = nj;
        if (sn) {
            sn[4 * j] = (float)nj.x;
            sn[4 * j + 1] = (float)nj.y;
            sn[4 * j + 2] = (float)nj.z;
            sn[4 * j + 3] = 0.0f;
        }
        if (L) {
            L[4 * j] = (float)Lj.x;
            L[4 * j + 1] = (float)Lj.y;
            L[4 * j + 2] = (float)Lj.z;
            L[4 * j + 3] = 0.0f;
        }
        if (w12) {
            w12[4 * j] = (float)wj.x;
            w12[4 * j + 1] = (float)wj.y;
            w12[4 * j + 2] = (float)wj.z;
            w12[4 * j + 3] = (float)Tj;
        }
    }
}

void orc_sss_coefficients(void* p, int m, float* out)
{
    const oscene* s = (const oscene*)p;
    const osss* q = &s->sss[m];
    const v3* v[] = {&q->absorption, &q->scattering, &q->meancosine, &q->extinction, &q->reducedExtinction,
                     &q->deltaEddExtinction, &q->D, &q->transport, &q->C_phi, &q->C_phi_inv, &q->C_E, &q->A,
                     &q->albedo, &q->reducedAlbedo, &q->de, &q->three_D, &q->two_A_de, &q->global_coeff,
                     &q->one_over_three_ext};
    int k = 0;
    out[k++] = (float)q->relative_ior;
    out[k++] = (float)q->dipole_model;
    for (int i = 0; i < 19; i++) {
        out[k++] = (float)v[i]->x;
        out[k++] = (float)v[i]->y;
        out[k++] = (float)v[i]->z;
    }
    out[k++] = (float)q->mean_transport;
    while (k < 64) out[k++] = 0.0f;
}

void orc_dipole_eval(void* p, int m, int n, const float* r, float* out)
{
    const oscene* s = (const oscene*)p;
    for (int i = 0; i < n; i++) {
        v3 v = dipole_bssrdf((R)r[i], &s->sss[m]);
        out[3 * i] = (float)v.x;
        out[3 * i + 1] = (float)v.y;
        out[3 * i + 2] = (float)v.z;
    }
}

/* S_infinite of material m for rows of 4 doubles (r^2, x.w12, no.w12, x.no),
 * the same value for the three channels' inputs; out: 3 doubles per row. */
void orc_dirpole_S_infinite_eval(void* p, int m, int n, const double* rows, double* out)
{
    const oscene* s = (const oscene*)p;
    for (int i = 0; i < n; i++) {
        const double* q = rows + 4 * i;
        R r2 = (R)q[0], xw = (R)q[1], xn = (R)q[3];
        v3 v = S_infinite(V3(r2, r2, r2), V3(xw, xw, xw), (R)q[2], V3(xn, xn, xn), &s->sss[m]);
        out[3 * i] = (double)v.x;
        out[3 * i + 1] = (double)v.y;
        out[3 * i + 2] = (double)v.z;
    }
}

void orc_dirpole_eval(void* p, int m, int n, const float* rows, float* out)
{
    const oscene* s = (const oscene*)p;
    for (int i = 0; i < n; i++) {
        const float* q = rows + 15 * i;
        v3 v = dirpole_bssrdf(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]), V3(q[9], q[10], q[11]),
                              V3(q[12], q[13], q[14]), &s->sss[m]);
        out[3 * i] = (float)v.x;
        out[3 * i + 1] = (float)v.y;
        out[3 * i + 2] = (float)v.z;
    }
}

/* Mean of `reps` gathers at (xo, no) on material m, gather seeds tea16(i, seed). */
void orc_sss_gather_eval(void* p, int m, const float* xo, const float* no, int reps, uint32_t seed, double* out)
{
    const oscene* s = (const oscene*)p;
    double acc[3] = {0, 0, 0};
    for (int i = 0; i < reps; i++) {
        uint32_t t64[2];
        t64[0] = t64[1] = orc_tea16((uint32_t)i, seed);
        v3 g = sss_gather(s, &s->sss[m], V3(xo[0], xo[1], xo[2]), V3(no[0], no[1], no[2]), t64);
        acc[0] += (double)g.x;
        acc[1] += (double)g.y;
        acc[2] += (double)g.z;
    }
    for (int k = 0; k < 3; k++) out[k] = acc[k] / reps;
}
