// orf_wavefront_host.h -- host-side driver of the batched wavefront integrator (internal).
#pragma once

#include "orf_devbuf.h"
#include "orf_wavefront.h"

namespace orf {

struct WavefrontConfig {
    int block_size = 128;         // workgroup size of the traversal kernels (64, 128, 256)
    int pool_size = 1 << 20;      // path slots in flight
    int sort_policy = 1;          // 1: per-material shade queues, 0: one shade kernel
    float regen_threshold = 1.0f; // refill free slots when continuing paths < threshold * pool
    int use_graph = 1;            // bounce iterations replayed as a hipGraph
    int graph_pairs = 4;          // iteration pairs per replay
};

struct WavefrontPool {
    DevBuf o, d, thr, L, st, hit, hleaf, flags, ext0, ext1, kind[Q_KINDS], sh_o, sh_d, sh_c, cnt, work;
    int cap = 0;
    hipStream_t stream = nullptr;
    unsigned* host_cnt = nullptr;
    long long last_iterations = 0;
    WavefrontPool() = default;
    WavefrontPool(const WavefrontPool&) = delete;
    WavefrontPool& operator=(const WavefrontPool&) = delete;
    ~WavefrontPool();
    void ensure(int capacity);
    WfView view(const SceneGPU& sc, const WfRender* renders, int nrenders, const int* remap, long long total,
                const WavefrontConfig& cfg, double* film, unsigned long long* stats);
};

void validate_config(const WavefrontConfig& c);

// Runs every work item of the batch; returns the wall time of the bounce loop (s).
double run_wavefront(WavefrontPool& pool, const SceneGPU& sc, const WfRender* d_renders, int nrenders,
                     const int* d_remap, long long total, const WavefrontConfig& cfg, double* film,
                     unsigned long long* stats, hipStream_t user_stream);

}  // namespace orf
