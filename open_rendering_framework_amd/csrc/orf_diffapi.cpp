// orf_diffapi.cpp -- C-ABI of the differentiable form (primal and adjoint launches).
#include <cstdlib>
#include <cstring>
#include <vector>

#include "orf_common.h"
#include "orf_context.h"
#include "orf_diff_host.h"

namespace orf {

static DiffParams diff_params(orf_context_t* c, const orf_diff_desc* d, const float* albedo, const float* rough,
                              const float* scattering)
{
    ORF_REQUIRE(d, "diff: null description");
    ORF_REQUIRE(c->analytic.empty(), "diff: scenes with disk or sphere lights are not supported");
    ORF_REQUIRE(d->tex_w > 0 && d->tex_h > 0 && d->tex_w <= 8192 && d->tex_h <= 8192, "diff: bad texture size");
    int nm = (int)c->mats.size();
    ORF_REQUIRE(d->albedo_material < nm && d->roughness_material < nm, "diff: material handle out of range");
    ORF_REQUIRE(d->albedo_material < 0 || albedo, "diff: albedo texture required");
    ORF_REQUIRE(d->roughness_material < 0 || rough, "diff: roughness texture required");
    ORF_REQUIRE(d->albedo_material < 0 || c->mats[d->albedo_material].kind == MAT_LAMBERT,
                "diff: the albedo texture applies to a Lambert (diffuse_shader) material");
    ORF_REQUIRE(d->roughness_material < 0 || (c->mats[d->roughness_material].kind == MAT_CONDUCTOR &&
                                              c->mats[d->roughness_material].normal_distribution == GGX_DISTRIBUTION),
                "diff: the roughness texture applies to a GGX conductor (metallic_shader) material");
    // d log f / d alpha of the single-scattering BSDF is the adjoint's roughness
    // term; the microsurface walk's weight is not that BSDF
    ORF_REQUIRE(d->roughness_material < 0 || c->mats[d->roughness_material].microfacet_model != MULTISCATTERING_MODEL,
                "diff: the roughness texture does not apply to the multiple-scattering (microsurface walk) model");
    ORF_REQUIRE(d->albedo_uv == 0 || d->albedo_uv == 1, "diff: albedo_uv must be 0 or 1");
    ORF_REQUIRE(d->roughness_uv == 0 || d->roughness_uv == 1, "diff: roughness_uv must be 0 or 1");
    DiffParams p;
    std::memset(&p, 0, sizeof(p));
    p.albedo = d->albedo_material >= 0 ? albedo : nullptr;
    p.rough = d->roughness_material >= 0 ? rough : nullptr;
    p.tw = d->tex_w;
    p.th = d->tex_h;
    p.albedo_mat = d->albedo_material;
    p.rough_mat = d->roughness_material;
    p.albedo_uv = d->albedo_uv;
    p.rough_uv = d->roughness_uv;
    for (int k = 0; k < 4; ++k) {
        p.albedo_uvp[k] = d->albedo_uv_params[k];
        p.rough_uvp[k] = d->roughness_uv_params[k];
    }
    p.sss_mat = -1;
    p.sss_scale = 1.0f;
    const int sm = d->scattering_material;
    ORF_REQUIRE(sm < nm, "diff: scattering material handle out of range");
    ORF_REQUIRE(sm < 0 || scattering, "diff: scattering coefficients required");
    if (sm >= 0) {
        ORF_REQUIRE(is_translucent(c->mats[sm].kind), "diff: the scattering coefficients apply to a translucent material");
        ORF_REQUIRE(!c->mat_sss[sm].has_user,
                    "diff: the scattering material takes its coefficients from user data, not from absorption/scattering");
        p.sss_mat = sm;
        p.sss_scale = c->mat_sss[sm].scale;
    }
    // gradient vertices an adjoint sample keeps before it replays its path
    // (kDiffRec; ORF_DIFF_RECORDS lowers it so the tests reach the replay)
    p.max_rec = kDiffRec;
    if (const char* e = std::getenv("ORF_DIFF_RECORDS")) {
        const int n = std::atoi(e);
        ORF_REQUIRE(n >= 0 && n <= kDiffRec, "diff: ORF_DIFF_RECORDS out of range");
        p.max_rec = n;
    }
    return p;
}

static void stage_views(orf_context_t* c, const orf_render_item* v, int n, hipStream_t st, int& w, int& h,
                        unsigned& spp)
{
    ORF_REQUIRE(v && n > 0 && n <= 4096, "diff: 1..4096 views required");
    w = v[0].width;
    h = v[0].height;
    spp = v[0].spp;
    std::vector<WfRender> R(n);
    for (int i = 0; i < n; ++i) {
        ORF_REQUIRE(v[i].width == w && v[i].height == h, "diff: all views must share one size");
        ORF_REQUIRE(v[i].spp == spp, "diff: all views must share one sample count");
        ORF_REQUIRE(w > 0 && h > 0 && w <= 16384 && h <= 16384 && v[i].spp > 0, "diff: bad view");
        std::memset(&R[i], 0, sizeof(WfRender));
        R[i].cam.eye = mk3(v[i].eye[0], v[i].eye[1], v[i].eye[2]);
        R[i].cam.U = mk3(v[i].U[0], v[i].U[1], v[i].U[2]);
        R[i].cam.V = mk3(v[i].V[0], v[i].V[1], v[i].V[2]);
        R[i].cam.W = mk3(v[i].W[0], v[i].W[1], v[i].W[2]);
        R[i].width = w;
        R[i].height = h;
        R[i].spp = v[i].spp;
        R[i].frame0 = v[i].frame0;
    }
    // the previous launch may still read the view table
    ORF_HIP_CHECK(hipStreamSynchronize(st));
    c->d_diff_views.upload(R);
    // [0] non-finite samples, [1] replay mismatches (adjoint), [8..15] work queues
    c->d_diff_stats.ensure(16 * sizeof(unsigned long long));
}

}  // namespace orf

namespace orf {

// The scene with the differentiated material's coefficients from `scattering`
// (device) in a copy of the coefficient table, and the probe pass run with it.
static SceneGPU diff_scene(orf_context_t* c, const DiffParams& dp, const float* scattering, unsigned frame,
                           hipStream_t st)
{
    SceneGPU sc = scene_view(c);
    ScatteringGPU* table = nullptr;
    if (dp.sss_mat >= 0) {
        c->d_diff_sss.ensure(c->d_sss.bytes);
        ORF_HIP_CHECK(hipMemcpyAsync(c->d_diff_sss.p, c->d_sss.p, c->d_sss.bytes, hipMemcpyDeviceToDevice, st));
        table = c->d_diff_sss.as<ScatteringGPU>();
        launch_sss_coeffs(scattering, dp.sss_scale, table, dp.sss_mat, st);
        sc.sss = table;
    }
    run_probe_pass(c, frame, st, table);
    return sc;
}

}  // namespace orf

namespace orf {

static void diff_render(orf_context_t* c, const orf_diff_desc* desc, const orf_render_item* views, int32_t nviews,
                        const float* albedo, const float* roughness, const float* scattering, float* images,
                        hipStream_t st)
{
    ORF_REQUIRE(images, "orf_diff_render: null image buffer");
    prepare(c, st);
    DiffParams dp = diff_params(c, desc, albedo, roughness, scattering);
    int w, h;
    unsigned spp;
    stage_views(c, views, nviews, st, w, h, spp);
    SceneGPU sc = diff_scene(c, dp, scattering, views[0].frame0, st);
    c->d_diff_partial.ensure((size_t)diff_partial_entries(nviews, w, h, spp) * sizeof(float4));
    unsigned long long* stats = c->d_diff_stats.as<unsigned long long>();
    launch_diff(sc, dp, c->d_diff_views.as<WfRender>(), nviews, w, h, spp, false, 0, images,
                c->d_diff_partial.as<float4>(), nullptr, nullptr, nullptr, nullptr, stats, stats + 8, st);
}

static void diff_adjoint(orf_context_t* c, const orf_diff_desc* desc, const orf_render_item* views, int32_t nviews,
                         const float* albedo, const float* roughness, const float* scattering, const float* d_images,
                         float* g_albedo, float* g_roughness, float* g_scattering, uint32_t frame_shift,
                         hipStream_t st)
{
    ORF_REQUIRE(d_images, "orf_diff_adjoint: null image gradient");
    prepare(c, st);
    DiffParams dp = diff_params(c, desc, albedo, roughness, scattering);
    int w, h;
    unsigned spp;
    stage_views(c, views, nviews, st, w, h, spp);
    // the probes of the primal's sample index, so the replay sees the primal's probe light
    SceneGPU sc = diff_scene(c, dp, scattering, views[0].frame0, st);
    unsigned long long* stats = c->d_diff_stats.as<unsigned long long>();
    launch_diff(sc, dp, c->d_diff_views.as<WfRender>(), nviews, w, h, spp, true, frame_shift, nullptr, nullptr,
                d_images, desc->albedo_material >= 0 ? g_albedo : nullptr,
                desc->roughness_material >= 0 ? g_roughness : nullptr, dp.sss_mat >= 0 ? g_scattering : nullptr, stats,
                stats + 8, st);
}

}  // namespace orf

using namespace orf;

extern "C" orf_status orf_diff_render(orf_context ctx, const orf_diff_desc* desc, const orf_render_item* views,
                                      int32_t nviews, const float* albedo, const float* roughness, float* images,
                                      void* stream)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(desc && desc->scattering_material < 0, "orf_diff_render: scattering coefficients need orf_diff_render_ex");
    diff_render(c, desc, views, nviews, albedo, roughness, nullptr, images, (hipStream_t)stream);
    ORF_API_END
}

extern "C" orf_status orf_diff_adjoint(orf_context ctx, const orf_diff_desc* desc, const orf_render_item* views,
                                       int32_t nviews, const float* albedo, const float* roughness,
                                       const float* d_images, float* g_albedo, float* g_roughness,
                                       uint32_t frame_shift, void* stream)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(desc && desc->scattering_material < 0,
                "orf_diff_adjoint: scattering coefficients need orf_diff_adjoint_ex");
    diff_adjoint(c, desc, views, nviews, albedo, roughness, nullptr, d_images, g_albedo, g_roughness, nullptr,
                 frame_shift, (hipStream_t)stream);
    ORF_API_END
}

extern "C" orf_status orf_diff_render_ex(orf_context ctx, const orf_diff_desc* desc, const orf_render_item* views,
                                         int32_t nviews, const float* albedo, const float* roughness,
                                         const float* scattering, float* images, void* stream)
{
    ORF_API_BEGIN
    diff_render(ctx_check(ctx), desc, views, nviews, albedo, roughness, scattering, images, (hipStream_t)stream);
    ORF_API_END
}

extern "C" orf_status orf_diff_adjoint_ex(orf_context ctx, const orf_diff_desc* desc, const orf_render_item* views,
                                          int32_t nviews, const float* albedo, const float* roughness,
                                          const float* scattering, const float* d_images, float* g_albedo,
                                          float* g_roughness, float* g_scattering, uint32_t frame_shift, void* stream)
{
    ORF_API_BEGIN
    ORF_REQUIRE(!desc || desc->scattering_material < 0 || g_scattering, "orf_diff_adjoint_ex: null scattering gradient");
    diff_adjoint(ctx_check(ctx), desc, views, nviews, albedo, roughness, scattering, d_images, g_albedo, g_roughness,
                 g_scattering, frame_shift, (hipStream_t)stream);
    ORF_API_END
}

extern "C" orf_status orf_diff_counters(orf_context ctx, uint64_t out[2])
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(out, "orf_diff_counters: null output");
    out[0] = out[1] = 0;
    if (!c->d_diff_stats.p) return ORF_SUCCESS;
    ORF_HIP_CHECK(hipDeviceSynchronize());
    ORF_HIP_CHECK(hipMemcpy(out, c->d_diff_stats.p, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    ORF_API_END
}
