// orf_batch.cpp -- C-ABI of batched renders: wavefront or per-item megakernel, autotuner.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "orf_common.h"
#include "orf_context.h"
#include "orf_mega.h"

namespace orf {

static WavefrontConfig to_cfg(const orf_wavefront_config* c)
{
    WavefrontConfig w;
    if (c) {
        w.block_size = c->block_size;
        w.pool_size = c->pool_size;
        w.sort_policy = c->sort_policy;
        w.regen_threshold = c->regen_threshold;
        w.use_graph = c->use_graph;
        w.graph_pairs = c->graph_pairs;
    }
    return w;
}

static void from_cfg(const WavefrontConfig& w, orf_wavefront_config* c)
{
    c->block_size = w.block_size;
    c->pool_size = w.pool_size;
    c->sort_policy = w.sort_policy;
    c->regen_threshold = w.regen_threshold;
    c->use_graph = w.use_graph;
    c->graph_pairs = w.graph_pairs;
}

// Build and upload the item table, remap table and a zeroed batch film.
static long long stage_batch(orf_context_t* c, const orf_render_item* items, int n, const int32_t* remap,
                             int stride, unsigned spp_div)
{
    ORF_REQUIRE(items && n > 0 && n <= 65536, "render batch: 1..65536 items required");
    int nm = (int)c->mats.size();
    ORF_REQUIRE(nm > 0, "render batch: no materials");
    std::vector<int> rm;
    if (remap) {
        ORF_REQUIRE(stride > 0, "render batch: remap_stride must be positive");
        rm.assign(remap, remap + (size_t)n * stride);
        for (int v : rm) ORF_REQUIRE(v >= 0 && v < nm, "render batch: remap entry is not a material handle");
    } else {
        stride = nm;
        for (int i = 0; i < n; ++i)
            for (int m = 0; m < nm; ++m) rm.push_back(m);
    }
    // geometry material ids must be valid indices into every item's table
    for (const MeshRec& m : c->meshes)
        ORF_REQUIRE(m.material < stride, "render batch: remap_stride smaller than a mesh material id");
    for (int m : c->analytic_mat)
        ORF_REQUIRE(m < stride, "render batch: remap_stride smaller than a disk/sphere emitter material id");
    c->batch_renders.clear();
    long long paths = 0, pixels = 0;
    for (int i = 0; i < n; ++i) {
        const orf_render_item& it = items[i];
        ORF_REQUIRE(it.width > 0 && it.height > 0 && it.width <= 16384 && it.height <= 16384,
                    "render batch: bad item size");
        WfRender r;
        std::memset(&r, 0, sizeof(r));
        r.cam.eye = mk3(it.eye[0], it.eye[1], it.eye[2]);
        r.cam.U = mk3(it.U[0], it.U[1], it.U[2]);
        r.cam.V = mk3(it.V[0], it.V[1], it.V[2]);
        r.cam.W = mk3(it.W[0], it.W[1], it.W[2]);
        r.cam.patch_dims[0] = (unsigned)it.width;
        r.cam.patch_dims[1] = (unsigned)it.height;
        r.width = it.width;
        r.height = it.height;
        r.frame0 = it.frame0;
        r.spp = std::max(1u, it.spp / std::max(1u, spp_div));
        if (it.spp == 0) r.spp = 0;
        r.path_begin = paths;
        r.film_offset = pixels;
        r.remap_offset = i * stride;
        paths += (long long)it.width * it.height * r.spp;
        pixels += (long long)it.width * it.height;
        c->batch_renders.push_back(r);
    }
    c->d_batch_renders.upload(c->batch_renders);
    c->d_batch_remap.upload(rm);
    c->d_batch_film.ensure((size_t)pixels * 4 * sizeof(double));
    ORF_HIP_CHECK(hipMemset(c->d_batch_film.p, 0, (size_t)pixels * 4 * sizeof(double)));
    c->d_batch_stats.ensure(16 * sizeof(unsigned long long));  // [0..3] counters, [8..15] queue counters
    ORF_HIP_CHECK(hipMemset(c->d_batch_stats.p, 0, c->d_batch_stats.bytes));
    c->batch_pixels = pixels;
    return paths;
}

static double run_batch(orf_context_t* c, long long paths, int mode, const WavefrontConfig& cfg, hipStream_t st)
{
    SceneGPU sc = scene_view(c);
    double* film = c->d_batch_film.as<double>();
    auto* stats = c->d_batch_stats.as<unsigned long long>();
    ORF_HIP_CHECK(hipDeviceSynchronize());
    auto t0 = std::chrono::steady_clock::now();
    run_probe_pass(c, c->batch_renders.empty() ? 0u : c->batch_renders[0].frame0, st);
    if (mode == ORF_BATCH_PERSISTENT) {
        // work items: (render, 8x8 tile, chunk of up to kMegaChunk samples), render by render
        std::vector<PersistRender> pr(c->batch_renders.size());
        long long items = 0;
        for (size_t i = 0; i < pr.size(); ++i) {
            const WfRender& r = c->batch_renders[i];
            PersistRender& p = pr[i];
            p.item_begin = items;
            p.tiles_x = (r.width + 7) / 8;
            p.per = std::max(1u, std::min(r.spp, kMegaChunk));
            p.chunks = (int)((r.spp + p.per - 1) / p.per);
            p.pad = 0;
            items += (long long)p.tiles_x * ((r.height + 7) / 8) * p.chunks;
        }
        c->d_batch_items.upload(pr);
        launch_batch_persistent(sc, c->d_batch_renders.as<WfRender>(), c->d_batch_items.as<PersistRender>(),
                                (int)pr.size(), c->d_batch_remap.as<int>(), film, stats, stats + 8, items,
                                persistent_waves(items, mega_waves(sc.features)), st);
        ORF_HIP_CHECK(hipStreamSynchronize(st));
    } else if (mode == ORF_BATCH_MEGAKERNEL) {
        const int* remap = c->d_batch_remap.as<int>();
        for (const WfRender& r : c->batch_renders) {
            FilmGPU f{film + 4 * (size_t)r.film_offset, stats, r.width, r.height, stats + 8};
            launch_megakernel(sc, r.cam, f, remap + r.remap_offset, r.frame0, r.spp, st);
        }
        ORF_HIP_CHECK(hipStreamSynchronize(st));
    } else {
        run_wavefront(c->wf_pool, sc, c->d_batch_renders.as<WfRender>(), (int)c->batch_renders.size(),
                      c->d_batch_remap.as<int>(), paths, cfg, film, stats, st);
        unsigned long long p = (unsigned long long)paths;
        ORF_HIP_CHECK(hipMemcpy(stats + 3, &p, 8, hipMemcpyHostToDevice));
        c->last_batch_iterations = c->wf_pool.last_iterations;
    }
    ORF_HIP_CHECK(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace orf

using namespace orf;

extern "C" orf_status orf_wavefront_default_config(orf_wavefront_config* cfg)
{
    ORF_API_BEGIN
    ORF_REQUIRE(cfg, "orf_wavefront_default_config: null output");
    from_cfg(WavefrontConfig(), cfg);
    ORF_API_END
}

extern "C" orf_status orf_render_batch(orf_context ctx, const orf_render_item* items, int32_t n,
                                       const int32_t* remap, int32_t remap_stride, int32_t mode,
                                       const orf_wavefront_config* cfg, void* stream, double* seconds)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(mode == ORF_BATCH_WAVEFRONT || mode == ORF_BATCH_MEGAKERNEL || mode == ORF_BATCH_PERSISTENT,
                "orf_render_batch: bad mode");
    WavefrontConfig w = to_cfg(cfg);
    if (mode == ORF_BATCH_WAVEFRONT) validate_config(w);
    hipStream_t st = (hipStream_t)stream;
    prepare(c, st);
    long long paths = stage_batch(c, items, n, remap, remap_stride, 1);
    double s = run_batch(c, paths, mode, w, st);
    c->last_batch_seconds = s;
    if (seconds) *seconds = s;
    ORF_API_END
}

extern "C" orf_status orf_batch_film_sums(orf_context ctx, int32_t item, double* sums)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(sums && item >= 0 && item < (int)c->batch_renders.size(), "orf_batch_film_sums: bad item");
    const WfRender& r = c->batch_renders[item];
    ORF_HIP_CHECK(hipDeviceSynchronize());
    ORF_HIP_CHECK(hipMemcpy(sums, c->d_batch_film.as<double>() + 4 * (size_t)r.film_offset,
                            (size_t)r.width * r.height * 4 * sizeof(double), hipMemcpyDeviceToHost));
    ORF_API_END
}

extern "C" orf_status orf_batch_film_read(orf_context ctx, int32_t item, double* rgba)
{
    ORF_API_BEGIN
    orf_status s = orf_batch_film_sums(ctx, item, rgba);
    if (s != ORF_SUCCESS) return s;
    const WfRender& r = ctx->batch_renders[item];
    size_t np = (size_t)r.width * r.height;
    for (size_t i = 0; i < np; ++i) {
        double cnt = rgba[4 * i + 3];
        for (int k = 0; k < 3; ++k) rgba[4 * i + k] = cnt > 0 ? rgba[4 * i + k] / cnt : 0.0;
        rgba[4 * i + 3] = 1.0;
    }
    ORF_API_END
}

extern "C" orf_status orf_batch_counters(orf_context ctx, uint64_t out[4])
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(out, "orf_batch_counters: null output");
    if (!c->d_batch_stats.p) {
        for (int k = 0; k < 4; ++k) out[k] = 0;
        return ORF_SUCCESS;
    }
    ORF_HIP_CHECK(hipDeviceSynchronize());
    ORF_HIP_CHECK(hipMemcpy(out, c->d_batch_stats.p, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    ORF_API_END
}

extern "C" orf_status orf_wavefront_autotune(orf_context ctx, const orf_render_item* items, int32_t n,
                                             const int32_t* remap, int32_t remap_stride, uint32_t spp_divisor,
                                             orf_wavefront_config* best, char* report, size_t report_size)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(best, "orf_wavefront_autotune: null output");
    prepare(c, nullptr);
    long long paths = stage_batch(c, items, n, remap, remap_stride, std::max(1u, spp_divisor));
    std::string rep = "{\"probe_paths\": " + std::to_string(paths) + ", \"trials\": [";
    bool first = true;
    auto trial = [&](const WavefrontConfig& w) {
        validate_config(w);
        ORF_HIP_CHECK(hipMemset(c->d_batch_film.p, 0, (size_t)c->batch_pixels * 4 * sizeof(double)));
        double s = run_batch(c, paths, ORF_BATCH_WAVEFRONT, w, nullptr);
        char buf[256];
        std::snprintf(buf, sizeof(buf),
                      "%s{\"block_size\": %d, \"pool_size\": %d, \"sort_policy\": %d, \"regen_threshold\": %.3f, "
                      "\"seconds\": %.6f, \"mpaths_per_s\": %.3f}",
                      first ? "" : ", ", w.block_size, w.pool_size, w.sort_policy, w.regen_threshold, s,
                      paths / s / 1e6);
        rep += buf;
        first = false;
        return s;
    };
    WavefrontConfig cur;
    double best_s = trial(cur);
    // coordinate search, one parameter at a time
    for (int b : {64, 256}) {
        WavefrontConfig w = cur;
        w.block_size = b;
        double s = trial(w);
        if (s < best_s) best_s = s, cur = w;
    }
    for (int p : {1 << 19, 1 << 21}) {
        WavefrontConfig w = cur;
        w.pool_size = p;
        double s = trial(w);
        if (s < best_s) best_s = s, cur = w;
    }
    {
        WavefrontConfig w = cur;
        w.sort_policy = 1 - cur.sort_policy;
        double s = trial(w);
        if (s < best_s) best_s = s, cur = w;
    }
    {
        WavefrontConfig w = cur;
        w.regen_threshold = 0.5f;
        double s = trial(w);
        if (s < best_s) best_s = s, cur = w;
    }
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  "], \"best\": {\"block_size\": %d, \"pool_size\": %d, \"sort_policy\": %d, \"regen_threshold\": %.3f, "
                  "\"seconds\": %.6f}}",
                  cur.block_size, cur.pool_size, cur.sort_policy, cur.regen_threshold, best_s);
    rep += buf;
    from_cfg(cur, best);
    if (report && report_size > 0) {
        std::strncpy(report, rep.c_str(), report_size - 1);
        report[report_size - 1] = 0;
    }
    ORF_API_END
}
