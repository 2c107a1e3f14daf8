// orf_wavefront.hip -- batched wavefront integrator kernels and their driver (gfx950).
//
// Stages and layout are described in orf_wavefront.h.  One bounce iteration is
//   reset(p) -> regen(p) -> extend(p) -> shade(kind...) -> shadow
// where p is the parity of the extension queue being consumed.  Two iterations
// (p = 0, 1) are captured into one hipGraph and replayed until every work item
// has been traced and flushed; all launch shapes are fixed (grid-stride loops
// over device-side queue counts), so replays need no host decisions.  Counters
// are reset by a one-lane kernel, never by memset nodes.  The traversal and
// shade kernels are compiled per material program set (kFeat, as the
// megakernel's variants): with per-material queues, the miss, emitter, Lambert
// and conductor queues run the smallest set whatever the scene holds.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "orf_common.h"
#include "orf_wavefront.h"
#include "orf_wavefront_host.h"

namespace orf {

// uniform trip count per workgroup: every lane of a wave runs the same iterations
#define WF_FOR(n, i, active)                                                                            \
    for (unsigned base_ = blockIdx.x * blockDim.x, end_ = (((unsigned)(n)) + blockDim.x - 1u) / blockDim.x * \
                                                          blockDim.x;                                   \
         base_ < end_; base_ += gridDim.x * blockDim.x)                                                  \
        for (unsigned i = base_ + threadIdx.x, once_ = 1; once_; once_ = 0)                              \
            for (bool active = i < (unsigned)(n); once_; once_ = 0)

__global__ void k_wf_reset(WfView v, int parity)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int k = 0; k < Q_KINDS; ++k) v.cnt[C_KIND0 + k] = 0;
    v.cnt[C_SHADOW] = 0;
    v.cnt[C_EXT0 + (parity ^ 1)] = 0;
    v.cnt[C_COUNT] = v.cnt[C_EXT0 + parity];  // continuing paths: snapshot for regen
}

__global__ __launch_bounds__(256) void k_wf_regen(WfView v, int parity)
{
    unsigned cont = v.cnt[C_COUNT];
    bool refill = (float)cont < v.regen_threshold * (float)v.capacity;
    unsigned* qcnt = v.cnt + C_EXT0 + parity;
    int* q = v.ext_q[parity];
    WF_FOR(v.capacity, i, active)
    {
        int f = active ? v.flags[i] : -1;
        if (f == SLOT_DONE) {
            flush_slot(v, (int)i);
            v.flags[i] = SLOT_FREE;
            f = SLOT_FREE;
        }
        bool want = active && refill && f == SLOT_FREE;
        unsigned long long mask = __ballot(want);
        long long id = -1;
        if (mask) {
            unsigned lane = __lane_id();
            int leader = __ffsll((long long)mask) - 1;
            unsigned long long base = 0;
            if ((int)lane == leader) base = atomicAdd(v.work_next, (unsigned long long)__popcll(mask));
            base = __shfl(base, leader);
            if (want) id = (long long)(base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull)));
        }
        bool start = id >= 0 && id < v.total_work;
        if (start) start_path(v, (int)i, id);
        unsigned idx = wave_push(qcnt, start);
        if (start) q[idx] = (int)i;
    }
}

template <int BLOCK, unsigned kFeat>
__global__ __launch_bounds__(BLOCK) void k_wf_extend(WfView v, int parity)
{
    __shared__ unsigned stack_lds[kLdsStack * BLOCK];
    unsigned* stk = stack_lds + threadIdx.x;
    unsigned n = v.cnt[C_EXT0 + parity];
    const int* q = v.ext_q[parity];
    unsigned rays = 0;
    WF_FOR(n, i, active)
    {
        int slot = active ? q[i] : -1;
        int kind = -1;
        if (active) {
            float4 o = v.o[slot], d = v.d[slot];
            Hit h;
            trace<false, kFeat == kFeatAll, (kFeat != 0u)>(v.sc, mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z),
                                                           v.sc.scene_epsilon, kRayMax, h, stk, BLOCK);
            v.hit[slot] = make_float4(h.t, h.b, h.g, __uint_as_float((unsigned)h.prim));
            v.hleaf[slot] = h.leaf;
            kind = v.sort_policy ? hit_kind(v, slot, h) : Q_ALL;
            rays++;
        }
        if (v.sort_policy) {
            for (int k = 0; k < Q_ALL; ++k) {
                unsigned idx = wave_push(v.cnt + C_KIND0 + k, kind == k);
                if (kind == k) v.kind_q[k][idx] = slot;
            }
        } else {
            unsigned idx = wave_push(v.cnt + C_KIND0 + Q_ALL, active);
            if (active) v.kind_q[Q_ALL][idx] = slot;
        }
    }
    if (rays) atomicAdd(v.stats + 1, (unsigned long long)rays);
}

template <unsigned kFeat>
__global__ __launch_bounds__(256) void k_wf_shade(WfView v, int parity, int kind)
{
    unsigned n = v.cnt[C_KIND0 + kind];
    const int* q = v.kind_q[kind];
    unsigned* next_cnt = v.cnt + C_EXT0 + (parity ^ 1);
    int* next_q = v.ext_q[parity ^ 1];
    WF_FOR(n, i, active)
    {
        int slot = active ? q[i] : -1;
        ShadowRay sr;
        sr.valid = false;
        bool more = false;
        if (active) more = shade_slot<kFeat>(v, slot, sr);
        unsigned e = wave_push(next_cnt, more);
        if (more) next_q[e] = slot;
        unsigned s = wave_push(v.cnt + C_SHADOW, sr.valid);
        if (sr.valid) {
            v.sh_o[s] = make_float4(sr.o.x, sr.o.y, sr.o.z, sr.tmax);
            v.sh_d[s] = make_float4(sr.d.x, sr.d.y, sr.d.z, __uint_as_float((unsigned)slot));
            v.sh_c[s] = make_float4(sr.contrib.x, sr.contrib.y, sr.contrib.z, (float)sr.emit);
        }
    }
}

template <int BLOCK, unsigned kFeat>
__global__ __launch_bounds__(BLOCK) void k_wf_shadow(WfView v)
{
    __shared__ unsigned stack_lds[kLdsStack * BLOCK];
    unsigned* stk = stack_lds + threadIdx.x;
    unsigned n = v.cnt[C_SHADOW];
    unsigned rays = 0;
    WF_FOR(n, i, active)
    {
        if (active) {
            float4 o = v.sh_o[i], d = v.sh_d[i];
            int slot = (int)__float_as_uint(d.w);
            Hit h;
            bool occ = trace<true, kFeat == kFeatAll, (kFeat != 0u)>(v.sc, mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z),
                                                                     v.sc.scene_epsilon, o.w, h, stk, BLOCK);
            if (!occ) {
                float4 c = v.sh_c[i];
                float4 l = v.L[slot];
                f3 L = mk3(l.x, l.y, l.z) + mk3(c.x, c.y, c.z);
                v.L[slot] = make_float4(L.x, L.y, L.z, 0.0f);
                if (c.w >= 0.0f) {  // the segment re-arms (or clears) the path's emission flag
                    uint4 st = v.st[slot];
                    st.y = (st.y & 0x7FFFFFFFu) | ((c.w > 0.5f ? 1u : 0u) << 31);
                    v.st[slot] = st;
                }
            }
            rays++;
        }
    }
    if (rays) atomicAdd(v.stats + 2, (unsigned long long)rays);
}

// ------------------------------------------------------------------ driver

static const int kLeanKinds[] = {Q_MISS, Q_EMIT, Q_LAMBERT, Q_COND};

template <int BLOCK, unsigned kFeat>
static void enqueue_iteration(const WfView& v, int parity, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, st, v, parity);
    int g256 = std::max(1, std::min(grid * BLOCK / 256, (v.capacity + 255) / 256));
    hipLaunchKernelGGL(k_wf_regen, dim3(g256), dim3(256), 0, st, v, parity);
    hipLaunchKernelGGL((k_wf_extend<BLOCK, kFeat>), dim3(grid), dim3(BLOCK), 0, st, v, parity);
    if (v.sort_policy) {
        // the smallest program set leaves out disk/sphere light sampling, which
        // only scenes of the all-feature set hold
        constexpr unsigned kLean = kFeat == kFeatAll ? kFeatAll : 0u;
        for (int k : kLeanKinds) hipLaunchKernelGGL(k_wf_shade<kLean>, dim3(g256), dim3(256), 0, st, v, parity, k);
        hipLaunchKernelGGL(k_wf_shade<kFeat>, dim3(g256), dim3(256), 0, st, v, parity, (int)Q_DIEL);
    } else {
        hipLaunchKernelGGL(k_wf_shade<kFeat>, dim3(g256), dim3(256), 0, st, v, parity, (int)Q_ALL);
    }
    hipLaunchKernelGGL((k_wf_shadow<BLOCK, kFeat>), dim3(grid), dim3(BLOCK), 0, st, v);
}

template <unsigned kFeat>
static void enqueue_iteration_feat(const WfView& v, int block, int parity, int grid, hipStream_t st)
{
    if (block == 64) enqueue_iteration<64, kFeat>(v, parity, grid, st);
    else if (block == 128) enqueue_iteration<128, kFeat>(v, parity, grid, st);
    else enqueue_iteration<256, kFeat>(v, parity, grid, st);
}

// The scene's material program set picks the kernel variants, as for the megakernel.
static void enqueue_pairs(const WfView& v, int block, int grid, int pairs, hipStream_t st)
{
    for (int k = 0; k < pairs; ++k)
        for (int p = 0; p < 2; ++p) {
            if (v.sc.features == 0u) enqueue_iteration_feat<0u>(v, block, p, grid, st);
            else if (v.sc.features == kFeatDielectric) enqueue_iteration_feat<kFeatDielectric>(v, block, p, grid, st);
            else enqueue_iteration_feat<kFeatAll>(v, block, p, grid, st);
        }
    ORF_HIP_CHECK(hipGetLastError());
}

void WavefrontPool::ensure(int capacity)
{
    if (capacity <= cap && o.p) return;
    size_t n = (size_t)capacity;
    o.ensure(n * 16);
    d.ensure(n * 16);
    thr.ensure(n * 16);
    L.ensure(n * 16);
    st.ensure(n * 16);
    hit.ensure(n * 16);
    hleaf.ensure(n * 4);
    flags.ensure(n * 4);
    ext0.ensure(n * 4);
    ext1.ensure(n * 4);
    for (auto& k : kind) k.ensure(n * 4);
    sh_o.ensure(n * 16);
    sh_d.ensure(n * 16);
    sh_c.ensure(n * 16);
    cnt.ensure(64 * 4);
    work.ensure(8);
    cap = capacity;
}

WfView WavefrontPool::view(const SceneGPU& sc, const WfRender* renders, int nrenders, const int* remap,
                           long long total, const WavefrontConfig& cfg, double* film, unsigned long long* stats)
{
    WfView v;
    std::memset(&v, 0, sizeof(v));
    v.sc = sc;
    v.renders = renders;
    v.nrenders = nrenders;
    v.remap = remap;
    v.total_work = total;
    v.capacity = cfg.pool_size;
    v.sort_policy = cfg.sort_policy;
    v.regen_threshold = cfg.regen_threshold;
    v.o = o.as<float4>();
    v.d = d.as<float4>();
    v.thr = thr.as<float4>();
    v.L = L.as<float4>();
    v.st = st.as<uint4>();
    v.hit = hit.as<float4>();
    v.hleaf = hleaf.as<int>();
    v.flags = flags.as<int>();
    v.ext_q[0] = ext0.as<int>();
    v.ext_q[1] = ext1.as<int>();
    for (int k = 0; k < Q_KINDS; ++k) v.kind_q[k] = kind[k].as<int>();
    v.sh_o = sh_o.as<float4>();
    v.sh_d = sh_d.as<float4>();
    v.sh_c = sh_c.as<float4>();
    v.cnt = cnt.as<unsigned>();
    v.work_next = work.as<unsigned long long>();
    v.film = film;
    v.stats = stats;
    return v;
}

void validate_config(const WavefrontConfig& c)
{
    ORF_REQUIRE(c.block_size == 64 || c.block_size == 128 || c.block_size == 256,
                "wavefront: block_size must be 64, 128 or 256");
    ORF_REQUIRE(c.pool_size >= 1024 && c.pool_size <= (1 << 24) && c.pool_size % 256 == 0,
                "wavefront: pool_size must be a multiple of 256 in [1024, 2^24]");
    ORF_REQUIRE(c.sort_policy == 0 || c.sort_policy == 1, "wavefront: sort_policy must be 0 or 1");
    ORF_REQUIRE(c.regen_threshold > 0.0f && c.regen_threshold <= 1.0f,
                "wavefront: regen_threshold must be in (0, 1]");
}

double run_wavefront(WavefrontPool& pool, const SceneGPU& sc, const WfRender* d_renders, int nrenders,
                     const int* d_remap, long long total, const WavefrontConfig& cfg, double* film,
                     unsigned long long* stats, hipStream_t user_stream)
{
    validate_config(cfg);
    pool.ensure(cfg.pool_size);
    if (!pool.stream) ORF_HIP_CHECK(hipStreamCreateWithFlags(&pool.stream, hipStreamNonBlocking));
    if (!pool.host_cnt) ORF_HIP_CHECK(hipHostMalloc((void**)&pool.host_cnt, 64, hipHostMallocDefault));
    hipStream_t st = pool.stream;
    // order after the caller's prior work
    hipEvent_t ev;
    ORF_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    ORF_HIP_CHECK(hipEventRecord(ev, user_stream));
    ORF_HIP_CHECK(hipStreamWaitEvent(st, ev, 0));
    ORF_HIP_CHECK(hipMemsetAsync(pool.flags.p, 0, (size_t)cfg.pool_size * 4, st));
    ORF_HIP_CHECK(hipMemsetAsync(pool.cnt.p, 0, pool.cnt.bytes, st));
    ORF_HIP_CHECK(hipMemsetAsync(pool.work.p, 0, 8, st));

    WfView v = pool.view(sc, d_renders, nrenders, d_remap, total, cfg, film, stats);
    int waves_per_block = cfg.block_size / 64;
    int grid = std::max(1, std::min((cfg.pool_size + cfg.block_size - 1) / cfg.block_size, 8192 / waves_per_block));

    auto t0 = std::chrono::steady_clock::now();
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int pairs = std::max(1, std::min(cfg.graph_pairs, 16));
    if (cfg.use_graph) {
        ORF_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue_pairs(v, cfg.block_size, grid, pairs, st);
        ORF_HIP_CHECK(hipStreamEndCapture(st, &graph));
        ORF_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    long long iters = 0;
    // a path needs at most max_depth + 3 iterations once started and a slot is
    // refilled at least every iteration once free: bound the loop generously
    long long max_iters =
        4 * ((total + cfg.pool_size - 1) / cfg.pool_size + 1) * (long long)(sc.max_depth + 4) + 256;
    for (;;) {
        if (exec) ORF_HIP_CHECK(hipGraphLaunch(exec, st));
        else enqueue_pairs(v, cfg.block_size, grid, pairs, st);
        iters += 2 * pairs;
        ORF_HIP_CHECK(hipMemcpyAsync(pool.host_cnt, pool.cnt.p, 8, hipMemcpyDeviceToHost, st));
        ORF_HIP_CHECK(hipMemcpyAsync(pool.host_cnt + 2, pool.work.p, 8, hipMemcpyDeviceToHost, st));
        ORF_HIP_CHECK(hipStreamSynchronize(st));
        unsigned continuing = pool.host_cnt[C_EXT0];
        unsigned long long claimed;
        std::memcpy(&claimed, pool.host_cnt + 2, 8);
        if (continuing == 0 && claimed >= (unsigned long long)total) break;
        ORF_REQUIRE(iters < max_iters, "wavefront: iteration bound exceeded");
    }
    // flush the slots finished in the last iteration (no work left to start)
    if (exec) ORF_HIP_CHECK(hipGraphLaunch(exec, st));
    else enqueue_pairs(v, cfg.block_size, grid, 1, st);
    ORF_HIP_CHECK(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    if (exec) ORF_HIP_CHECK(hipGraphExecDestroy(exec));
    if (graph) ORF_HIP_CHECK(hipGraphDestroy(graph));
    ORF_HIP_CHECK(hipEventRecord(ev, st));
    ORF_HIP_CHECK(hipStreamWaitEvent(user_stream, ev, 0));
    ORF_HIP_CHECK(hipEventDestroy(ev));
    pool.last_iterations = iters;
    return std::chrono::duration<double>(t1 - t0).count();
}

WavefrontPool::~WavefrontPool()
{
    if (stream) (void)hipStreamDestroy(stream);
    if (host_cnt) (void)hipHostFree(host_cnt);
}

}  // namespace orf
