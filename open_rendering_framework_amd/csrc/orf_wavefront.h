// orf_wavefront.h -- batched wavefront integrator: data layout and per-item stages.
//
// Many renders share one scene; every (render, pixel, sample) path is one work
// item of a global list.  A fixed pool of path slots holds the paths in flight
// as structures of arrays; each bounce iteration runs
//   regen   : flush finished slots into their render's FP64 film, refill free
//             slots with new work items (camera rays) while work remains;
//   extend  : closest hit for every queued ray, then the slot is pushed into
//             the queue of its hit's material kind (or the single shade queue);
//   shade   : shade_vertex()/shade_miss() -- the same code the megakernel runs --
//             pushes continuing paths into the next extension queue and NEE
//             segments into the shadow queue;
//   shadow  : any hit for each shadow segment; unoccluded ones add to L.
// Queue pushes are wave-aggregated: one ballot, one atomic per wave and kind,
// ranks from a popcount of the lower lanes.  The per-item functions below are
// host-callable so the whole pipeline can be stepped on the CPU.
#pragma once

#include "orf_path.h"

namespace orf {

enum WfQueue : int { Q_MISS = 0, Q_EMIT = 1, Q_LAMBERT = 2, Q_COND = 3, Q_DIEL = 4, Q_ALL = 5, Q_KINDS = 6 };
enum WfSlot : int { SLOT_FREE = 0, SLOT_ACTIVE = 1, SLOT_DONE = 2 };
// counter slots
enum WfCounter : int { C_EXT0 = 0, C_EXT1 = 1, C_KIND0 = 2, C_SHADOW = C_KIND0 + Q_KINDS, C_COUNT = C_SHADOW + 1 };

struct WfRender {
    CameraGPU cam;
    int width, height;
    unsigned frame0, spp;
    long long path_begin;   // first global work item of this render
    long long film_offset;  // first pixel of this render in the batch film
    int remap_offset;       // first entry of this render's material table
    int pad;
};

struct WfView {
    SceneGPU sc;
    const WfRender* renders;
    int nrenders;
    const int* remap;       // per-render material tables
    long long total_work;
    int capacity;           // path slots
    int sort_policy;        // 0: one shade queue, 1: per-material queues
    float regen_threshold;  // refill when continuing paths < threshold * capacity
    // path pool (SoA, `capacity` entries)
    float4* o;        // origin xyz, w = bits of the 64-bit random state's first word
    float4* d;        // direction xyz, w = prev_bsdf_pdf
    float4* thr;      // throughput xyz, w = bits of the 64-bit random state's second word
    float4* L;        // radiance xyz
    uint4* st;        // rng, depth | emit << 31, render, pixel
    float4* hit;      // t, b, g, prim bits
    int* hleaf;       // leaf triangle of the hit
    int* flags;       // WfSlot
    // queues
    int* ext_q[2];
    int* kind_q[Q_KINDS];
    float4* sh_o;     // origin xyz, tmax
    float4* sh_d;     // direction xyz, slot bits
    float4* sh_c;     // contribution xyz
    // counters
    unsigned* cnt;                  // WfCounter
    unsigned long long* work_next;  // next unclaimed work item
    double* film;                   // 4 doubles per pixel of every render
    unsigned long long* stats;      // [0] non-finite, [1] ext rays, [2] shadow rays, [3] paths
};

// ------------------------------------------------------------ queue push

#if defined(__HIP_DEVICE_COMPILE__)
// All lanes of the wave must call this together (uniform trip counts).
__device__ __forceinline__ unsigned wave_push(unsigned* counter, bool pred)
{
    unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return ~0u;
    unsigned lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    unsigned base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (unsigned)__popcll(mask));
    base = __shfl(base, leader);
    unsigned rank = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
    return pred ? base + rank : ~0u;
}
#else
ORF_HD unsigned wave_push(unsigned* counter, bool pred) { return pred ? (*counter)++ : ~0u; }
#endif

ORF_HD void load_path(const WfView& v, int slot, PathState& ps)
{
    float4 o = v.o[slot], d = v.d[slot], t = v.thr[slot], l = v.L[slot];
    uint4 s = v.st[slot];
    ps.o = mk3(o.x, o.y, o.z);
    ps.d = mk3(d.x, d.y, d.z);
    ps.prev_bsdf_pdf = d.w;
    ps.thr = mk3(t.x, t.y, t.z);
    ps.L = mk3(l.x, l.y, l.z);
    ps.rng = s.x;
    ps.s64 = Seed64{f2u(o.w), f2u(t.w)};
    ps.depth = (int)(s.y & 0x7FFFFFFFu);
    ps.emit = (int)(s.y >> 31);
}

ORF_HD void store_path(const WfView& v, int slot, const PathState& ps)
{
    v.o[slot] = make_float4(ps.o.x, ps.o.y, ps.o.z, u2f(ps.s64.x));
    v.d[slot] = make_float4(ps.d.x, ps.d.y, ps.d.z, ps.prev_bsdf_pdf);
    v.thr[slot] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, u2f(ps.s64.y));
    v.L[slot] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
    uint4 s = v.st[slot];
    s.x = ps.rng;
    s.y = (unsigned)ps.depth | ((unsigned)(ps.emit ? 1 : 0) << 31);
    v.st[slot] = s;
}

ORF_HD int find_render(const WfView& v, long long id)
{
    int lo = 0, hi = v.nrenders - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (v.renders[mid].path_begin <= id) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Flush a finished slot into its film pixel.
ORF_HD void flush_slot(const WfView& v, int slot)
{
    uint4 s = v.st[slot];
    float4 l = v.L[slot];
    const WfRender& r = v.renders[s.z];
    double* px = v.film + 4 * (size_t)(r.film_offset + (long long)s.w);
    if (__builtin_isfinite(l.x) && __builtin_isfinite(l.y) && __builtin_isfinite(l.z)) {
#if defined(__HIP_DEVICE_COMPILE__)
        atomicAdd(px + 0, (double)l.x);
        atomicAdd(px + 1, (double)l.y);
        atomicAdd(px + 2, (double)l.z);
        atomicAdd(px + 3, 1.0);
#else
        px[0] += (double)l.x;
        px[1] += (double)l.y;
        px[2] += (double)l.z;
        px[3] += 1.0;
#endif
    } else {
#if defined(__HIP_DEVICE_COMPILE__)
        atomicAdd(v.stats + 0, 1ull);
#else
        v.stats[0] += 1;
#endif
    }
}

// Start work item `id` in `slot` (CU/path_tracer.cu:45-60 for sample frame0 + s).
ORF_HD void start_path(const WfView& v, int slot, long long id)
{
    int r = find_render(v, id);
    const WfRender& R = v.renders[r];
    long long local = id - R.path_begin;
    long long npix = (long long)R.width * R.height;
    unsigned s = (unsigned)(local / npix);
    int pix = (int)(local - (long long)s * npix);
    int x = pix % R.width, y = pix / R.width;
    f3 dir;
    uint32_t s0;
    uint32_t rng = camera_ray(R.cam, R.width, R.height, x, y, R.frame0 + s, dir, &s0);
    PathState ps;
    path_begin(ps, R.cam.eye, dir, rng, seed64(s0));
    uint4 st;
    st.x = 0;
    st.y = 0;
    st.z = (unsigned)r;
    st.w = (unsigned)pix;
    v.st[slot] = st;
    store_path(v, slot, ps);
    v.flags[slot] = SLOT_ACTIVE;
}

// Shade queue of the queued hit of `slot`.  Q_MISS, Q_EMIT, Q_LAMBERT and Q_COND
// hold only what the smallest material program set shades (misses, triangle
// emitters, Lambert, single-scattering conductors), so their shade kernels are
// compiled without the other materials (unless the scene needs the all-feature
// set for its disk or sphere lights); every other hit -- the remaining
// materials, multiple-scattering conductors, disk and sphere emitters -- goes
// to Q_DIEL, shaded by the scene's full program set.
ORF_HD int hit_kind(const WfView& v, int slot, const Hit& h)
{
    if (h.prim < 0) return Q_MISS;
    if (h.leaf < 0) return Q_DIEL;
    int4 sh = v.sc.shade[h.prim];
    const int* remap = v.remap + v.renders[v.st[slot].z].remap_offset;
    const MaterialGPU m = hit_material(v.sc, remap, sh.w);
    if (m.kind == MAT_EMITTER) return Q_EMIT;
    if (m.kind == MAT_LAMBERT) return Q_LAMBERT;
    if (m.kind == MAT_CONDUCTOR && m.microfacet_model != MULTISCATTERING_MODEL) return Q_COND;
    return Q_DIEL;
}

// Shade the hit (or miss) of `slot` with the material programs of kFeat; returns
// true when the path continues, and fills `sr` with an NEE segment when one is
// due.
template <unsigned kFeat>
ORF_HD bool shade_slot(const WfView& v, int slot, ShadowRay& sr)
{
    PathState ps;
    load_path(v, slot, ps);
    float4 hv = v.hit[slot];
    Hit h;
    h.t = hv.x;
    h.b = hv.y;
    h.g = hv.z;
    h.prim = (int)f2u(hv.w);
    h.leaf = v.hleaf[slot];
    sr.valid = false;
    bool more;
    if (h.prim < 0) {
        shade_miss(v.sc, ps);
        more = false;
    } else {
        const int* remap = v.remap + v.renders[v.st[slot].z].remap_offset;
        more = shade_vertex<kFeat>(v.sc, remap, ps, h, sr);
    }
    store_path(v, slot, ps);
    if (!more) v.flags[slot] = SLOT_DONE;
    return more;
}

}  // namespace orf
