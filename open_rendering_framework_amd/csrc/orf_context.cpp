// orf_context.cpp -- C-ABI render context: scene objects, named variables, BVH
// build and upload, launches and film I/O.
//
// The context plays the role of optix::Context plus the scene graph the
// reference's loader builds (ORF/OptixScene.cpp, ORF/OptixSceneLoader.cpp):
// host-side records are flattened to world space at build time (instance
// transforms applied to vertices, inverse-transpose to normals), one BVH is
// built over all triangles, and the device copies are refreshed only for what
// changed since the last launch.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "orf_common.h"
#include "orf_context.h"
#include "orf_math.h"

namespace orf {

}  // namespace orf

namespace orf {

orf_context_t* ctx_check(orf_context ctx)
{
    ORF_REQUIRE(ctx != nullptr, "null context");
    ORF_HIP_CHECK(hipSetDevice(ctx->device));
    return ctx;
}

MaterialGPU default_material(int kind)
{
    MaterialGPU m;
    std::memset(&m, 0, sizeof(m));
    m.kind = kind;
    m.microfacet_model = VISIBLE_NORMALS_MODEL;
    m.normal_distribution = GGX_DISTRIBUTION;
    m.diffuse_color = mk3(0.8f);
    m.radiance = mk3(1.0f);
    m.roughness = mk2(0.1f, 0.1f);
    m.ior = 1.5f;
    double re[3] = {1.5, 1.5, 1.5}, im[3] = {0.0, 0.0, 0.0};
    conductor_eta_host(re, im, m.eta);
    return m;
}

int program_kind(const std::string& prog)
{
    std::string p = prog;
    if (p.size() > 3 && p.compare(p.size() - 3, 3, ".cu") == 0) p = p.substr(0, p.size() - 3);
    if (p == "diffuse_shader") return MAT_LAMBERT;
    if (p == "metallic_shader") return MAT_CONDUCTOR;
    if (p == "arealight_shader") return MAT_EMITTER;
    if (p == "rough_transparent_shader") return MAT_DIELECTRIC;
    if (p == "lambertian_interface_shader") return MAT_INTERFACE;
    if (p == "subsurface_scattering_shader" || p == "translucent_shader") return MAT_TRANSLUCENT;
    if (p == "rough_subsurface_scattering_shader") return MAT_ROUGH_TRANSLUCENT;
    if (p == "transparent_shader") return MAT_TRANSPARENT;
    if (p == "normal_shader") return MAT_NORMAL;
    if (p == "flat_shader") return MAT_FLAT;
    if (p == "rough_diffuse_shader") return MAT_ROUGH_DIFFUSE;
    return -1;
}

static void set_context_float(orf_context_t* c, const std::string& name, const float* v, int n)
{
    auto need = [&](int k) { ORF_REQUIRE(n == k, "context variable '" + name + "' takes " + std::to_string(k) + " floats"); };
    if (name == "scene_epsilon") {
        need(1);
        c->scene_epsilon = v[0];
    } else if (name == "eye" || name == "U" || name == "V" || name == "W") {
        need(3);
        f3 x = mk3(v[0], v[1], v[2]);
        if (name == "eye") c->cam.eye = x;
        else if (name == "U") c->cam.U = x;
        else if (name == "V") c->cam.V = x;
        else c->cam.W = x;
    } else if (name == "exception_color") {
        need(3);
        for (int k = 0; k < 3; ++k) c->exception_color[k] = v[k];
    } else {
        ORF_THROW(ORF_ERROR_INVALID_VALUE, "unknown float context variable '" + name + "'");
    }
}

static void set_context_int(orf_context_t* c, const std::string& name, const int64_t* v, int n)
{
    if (name == "max_depth") {
        ORF_REQUIRE(n == 1 && v[0] >= 0, "max_depth takes one non-negative int");
        c->max_depth = (int)v[0];
    } else if (name == "frame") {
        ORF_REQUIRE(n == 1, "frame takes one value");
        c->frame = (unsigned)v[0];
    } else if (name == "patch_origin" || name == "patch_dims") {
        ORF_REQUIRE(n == 2, name + " takes two values");
        unsigned* dst = name == "patch_origin" ? c->cam.patch_origin : c->cam.patch_dims;
        dst[0] = (unsigned)v[0];
        dst[1] = (unsigned)v[1];
        c->patch_set = true;
    } else if (name == "envmap_importance") {
        // importance sampling of an environment bound by orf_variable_set_texture("envmap")
        ORF_REQUIRE(n == 1, "envmap_importance takes one value");
        c->envmap_importance = v[0] != 0;
    } else if (name == "probe_count" || name == "probe_samples" || name == "probe_seed") {
        // surface probes of the translucent object (count, gather samples per hit, seed)
        ORF_REQUIRE(n == 1 && v[0] >= 0 && v[0] <= (1 << 24), name + " takes one value in [0, 2^24]");
        if (name == "probe_count") c->probe_count_req = (int)v[0];
        else if (name == "probe_samples") c->probe_samples = (int)v[0];
        else c->probe_seed = (unsigned)v[0];
        if (name != "probe_samples") c->dirty_geo = true;
    } else {
        ORF_THROW(ORF_ERROR_INVALID_VALUE, "unknown integer context variable '" + name + "'");
    }
}

MaterialGPU& material_ref(orf_context_t* c, int handle)
{
    ORF_REQUIRE(handle >= 0 && handle < (int)c->mats.size(), "bad material handle");
    c->dirty_mat = true;
    return c->mats[handle];
}

void set_material_float(orf_context_t* c, int handle, const std::string& name, const float* v, int n)
{
    MaterialGPU& m = material_ref(c, handle);
    if (name == "diffuse_color" || name == "radiance" || name == "glass_absorption" || name == "highlight_color") {
        ORF_REQUIRE(n == 3, "material variable '" + name + "' takes 3 floats");
        f3 x = mk3(v[0], v[1], v[2]);
        if (name == "diffuse_color") m.diffuse_color = x;
        else if (name == "radiance" || name == "highlight_color") m.radiance = x;
        else m.glass_absorption = x;
    } else if (name == "highlight_threshold") {
        ORF_REQUIRE(n == 1, "highlight_threshold takes 1 float");
        m.threshold = v[0];
    } else if (name == "shininess") {
        ORF_REQUIRE(n == 1, "shininess takes 1 float");  // declared by FlatMaterial, unused by flat_shader.cu
    } else if (name == "roughness") {
        ORF_REQUIRE(n == 1 || n == 2, "roughness takes 1 or 2 floats");
        m.roughness = mk2(v[0], n == 2 ? v[1] : v[0]);
    } else if (name == "ior") {
        ORF_REQUIRE(n == 1, "float ior takes 1 float (conductors take MyComplex3 user data)");
        m.ior = v[0];
        c->mat_sss[handle].ior = v[0];
    } else if (name == "absorption" || name == "scattering" || name == "meancosine") {
        // ScatteringMaterial / TranslucentMaterial parameters (ORF/RoughTranslucentMaterial.cpp:158-215)
        ORF_REQUIRE(n == 3, "material variable '" + name + "' takes 3 floats");
        f3 x = mk3(v[0], v[1], v[2]);
        SssParamsHost& h = c->mat_sss[handle];
        if (name == "absorption") h.absorption = x;
        else if (name == "scattering") h.scattering = x;
        else h.meancosine = x;
    } else if (name == "scale") {
        ORF_REQUIRE(n == 1 && v[0] > 0.0f, "scale takes one positive float");
        c->mat_sss[handle].scale = v[0];
    } else {
        ORF_THROW(ORF_ERROR_INVALID_VALUE, "unknown float material variable '" + name + "'");
    }
}

void set_material_int(orf_context_t* c, int handle, const std::string& name, const int64_t* v, int n)
{
    MaterialGPU& m = material_ref(c, handle);
    ORF_REQUIRE(n == 1, "material variable '" + name + "' takes one value");
    if (name == "microfacet_model") {
        ORF_REQUIRE(v[0] >= 0 && v[0] <= 2, "microfacet_model must be 0, 1 or 2");
        m.microfacet_model = (int)v[0];
    } else if (name == "normal_distribution") {
        ORF_REQUIRE(v[0] == 0 || v[0] == 1, "normal_distribution must be 0 or 1");
        m.normal_distribution = (int)v[0];
    } else if (name == "dipole_model") {
        ORF_REQUIRE(v[0] == 0 || v[0] == 1, "dipole_model must be 0 (standard) or 1 (directional)");
        c->mat_sss[handle].dipole_model = (int)v[0];
    } else if (name == "flat_shadow" || name == "highlight") {
        const int bit = name == "flat_shadow" ? kFlatShadow : kFlatHighlight;
        m.flags = v[0] == 1 ? (m.flags | bit) : (m.flags & ~bit);  // the shader tests == 1
    } else {
        ORF_THROW(ORF_ERROR_INVALID_VALUE, "unknown integer material variable '" + name + "'");
    }
}

// Flatten every mesh to world space, build the BVH, upload geometry tables.
static void build_geometry(orf_context_t* c, hipStream_t)
{
    std::vector<float> V, N;
    std::vector<int32_t> I;
    std::vector<int4> shade;
    for (const MeshRec& m : c->meshes) {
        int32_t vbase = (int32_t)(V.size() / 3), nbase = (int32_t)(N.size() / 3);
        double M[16], Nm[9];
        if (m.has_xf) {
            for (int k = 0; k < 16; ++k) M[k] = m.xf[k];
            // inverse transpose of the upper 3x3 (cofactor matrix / det)
            double a = M[0], b = M[1], cc = M[2], d = M[4], e = M[5], f = M[6], g = M[8], h = M[9], i = M[10];
            double det = a * (e * i - f * h) - b * (d * i - f * g) + cc * (d * h - e * g);
            ORF_REQUIRE(det != 0.0, "singular mesh transform");
            double cof[9] = {e * i - f * h, -(d * i - f * g), d * h - e * g, -(b * i - cc * h), a * i - cc * g,
                             -(a * h - b * g), b * f - cc * e, -(a * f - cc * d), a * e - b * d};
            for (int k = 0; k < 9; ++k) Nm[k] = cof[k] / det;
        }
        for (size_t k = 0; k < m.v.size(); k += 3) {
            if (m.has_xf) {
                double p[3] = {m.v[k], m.v[k + 1], m.v[k + 2]};
                for (int r = 0; r < 3; ++r)
                    V.push_back((float)(M[4 * r] * p[0] + M[4 * r + 1] * p[1] + M[4 * r + 2] * p[2] + M[4 * r + 3]));
            } else {
                V.insert(V.end(), {m.v[k], m.v[k + 1], m.v[k + 2]});
            }
        }
        for (size_t k = 0; k < m.n.size(); k += 3) {
            if (m.has_xf) {
                double p[3] = {m.n[k], m.n[k + 1], m.n[k + 2]};
                for (int r = 0; r < 3; ++r)
                    N.push_back((float)(Nm[3 * r] * p[0] + Nm[3 * r + 1] * p[1] + Nm[3 * r + 2] * p[2]));
            } else {
                N.insert(N.end(), {m.n[k], m.n[k + 1], m.n[k + 2]});
            }
        }
        size_t nt = m.vi.size() / 3;
        for (size_t t = 0; t < nt; ++t) {
            for (int k = 0; k < 3; ++k) I.push_back(m.vi[3 * t + k] + vbase);
            int4 s;
            if (m.ni.empty()) {
                s = make_int4(-1, -1, -1, m.material);
            } else {
                s = make_int4(m.ni[3 * t] + nbase, m.ni[3 * t + 1] + nbase, m.ni[3 * t + 2] + nbase, m.material);
            }
            shade.push_back(s);
        }
    }
    // disk / sphere emitters: shade records after the triangles'
    c->analytic_prim = (int)shade.size();
    for (int m : c->analytic_mat) shade.push_back(make_int4(-1, -1, -1, m));
    BvhResult bvh = build_bvh(V.data(), (int64_t)(V.size() / 3), I.data(), (int64_t)(I.size() / 3));
    c->bvh_stats = bvh.stats;
    // triangles whose material lets shadow rays through
    for (size_t k = 0; k < bvh.tris.size(); k += 3) {
        uint32_t id;
        std::memcpy(&id, &bvh.tris[k].w, 4);
        const int m = shade[id].w;
        if (m >= 0 && m < (int)c->mats.size() && c->mats[m].kind == MAT_NORMAL) {
            ORF_REQUIRE(id <= kTriIdMask, "too many triangles");
            id |= kTriNoShadow;
            std::memcpy(&bvh.tris[k].w, &id, 4);
        }
    }
    if (bvh.tris.empty()) bvh.tris.assign(3, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    if (shade.empty()) shade.push_back(make_int4(-1, -1, -1, 0));
    std::vector<float4> normals(N.size() / 3 + 1);
    for (size_t k = 0; k + 2 < N.size(); k += 3) normals[k / 3] = make_float4(N[k], N[k + 1], N[k + 2], 0.0f);
    ORF_REQUIRE(bvh.stack4 <= kStackDepth, "scene: the BVH needs a deeper traversal stack than this build has");
    c->d_bvh4.upload(bvh.nodes4);
    c->nodes4 = (int)(bvh.nodes4.size() / 8);
    c->d_tri.upload(bvh.tris);
    c->d_shade.upload(shade);
    c->d_normals.upload(normals);
    std::vector<float4> an = c->analytic;
    if (an.empty()) an.assign(2, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    c->d_analytic.upload(an);
    // probes on the first mesh whose material is translucent
    c->probe_sss = -1;
    c->probe_count = 0;
    if (c->probe_count_req > 0) {
        for (const MeshRec& m : c->meshes)
            if (m.material >= 0 && m.material < (int)c->mats.size() && is_translucent(c->mats[m.material].kind)) {
                c->probe_sss = m.material;
                break;
            }
    }
    if (c->probe_sss >= 0) {
        ProbeSetHost ps = generate_probes(V.data(), I.data(), shade.data(), normals.data(), (int64_t)(I.size() / 3),
                                          c->probe_sss, c->probe_count_req, c->probe_seed);
        c->probe_count = (int)ps.pos.size();
        c->probe_mask = ps.table_mask;
        c->probe_cell = ps.cell;
        c->probe_rmin = ps.rmin;
        c->probe_area = ps.area;
        if (c->probe_count > 0) {
            c->d_probe_pos.upload(ps.pos);
            c->d_probe_nrm.upload(ps.nrm);
            c->d_probe_cells.upload(ps.cells);
            c->d_probe_L.ensure(ps.pos.size() * sizeof(float4));
            c->d_probe_w12.ensure(ps.pos.size() * sizeof(float4));
            c->d_probe_sn.ensure(ps.pos.size() * sizeof(float4));
            ORF_HIP_CHECK(hipMemset(c->d_probe_L.p, 0, c->d_probe_L.bytes));
            ORF_HIP_CHECK(hipMemset(c->d_probe_w12.p, 0, c->d_probe_w12.bytes));
            ORF_HIP_CHECK(hipMemset(c->d_probe_sn.p, 0, c->d_probe_sn.bytes));
        } else {
            c->probe_sss = -1;
        }
    }
    c->dirty_geo = false;
}

static void upload_env(orf_context_t* c, hipStream_t stream)
{
    if (c->env_w > 0) {
        c->d_env.upload(c->env_tex);
        size_t w = c->env_w, h = c->env_h;
        c->d_mpdf.ensure(h * 4);
        c->d_mcdf.ensure(h * 4);
        c->d_cpdf.ensure(w * h * 4);
        c->d_ccdf.ensure(w * h * 4);
        c->d_envscratch.ensure((h + 1) * 8);
        if (c->env_is) {
            build_env_tables(c->d_env.as<float4>(), c->env_w, c->env_h, c->d_mpdf.as<float>(), c->d_mcdf.as<float>(),
                             c->d_cpdf.as<float>(), c->d_ccdf.as<float>(), c->d_envscratch.as<double>(), stream);
            ORF_HIP_CHECK(hipStreamSynchronize(stream));
        }
    }
    c->dirty_env = false;
    c->dirty_light = true;  // the environment entry of the light list follows env_is
}

static void upload_lights(orf_context_t* c)
{
    std::vector<LightGPU> L = c->lights;
    for (size_t i = 0; i < L.size(); ++i) {
        if (L[i].type != TRIANGLES_AREA_LIGHT) continue;
        // {count, first, last} into triangle_light_buffer (ORF/OptixSceneLoader.cpp:305-317)
        const float cnt = L[i].params.x, first = L[i].params.y, last = L[i].params.z;
        ORF_REQUIRE(cnt >= 1.0f && first >= 0.0f && last == first + cnt - 1.0f &&
                        last < (float)c->tri_lights.size() && cnt == std::floor(cnt) && first == std::floor(first),
                    "light " + std::to_string(i) + ": its triangles lie outside triangle_light_buffer");
    }
    if (c->env_w > 0 && c->env_is) {
        LightGPU e;
        std::memset(&e, 0, sizeof(e));
        e.type = ENV_LIGHT;
        L.push_back(e);
    }
    c->nlights_dev = (int)L.size();
    if (L.empty()) L.resize(1);
    c->d_lights.upload(L);
    std::vector<TriLightGPU> T = c->tri_lights;
    if (T.empty()) T.resize(1);
    c->d_trilights.upload(T);
    c->dirty_light = false;
}

void prepare(orf_context_t* c, hipStream_t stream)
{
    if (c->dirty_geo) build_geometry(c, stream);
    if (c->dirty_mat) {
        std::vector<MaterialGPU> m = c->mats;
        std::vector<ScatteringGPU> sss;
        for (size_t i = 0; i < m.size(); ++i) {
            sss.push_back(scattering_coefficients(c->mat_sss[i]));
            m[i].sss = is_translucent(m[i].kind) ? (int)i : -1;
            // the rough diffuse interface reflects as a conductor of index 1 / ior
            if (m[i].kind == MAT_ROUGH_DIFFUSE) dielectric_eta_host(1.0f / m[i].ior, m[i].eta);
        }
        if (m.empty()) m.push_back(default_material(MAT_LAMBERT));
        if (sss.empty()) sss.push_back(scattering_coefficients(SssParamsHost()));
        c->d_mats.upload(m);
        c->d_sss.upload(sss);
        c->dirty_mat = false;
    }
    if (c->dirty_env) upload_env(c, stream);
    if (c->dirty_light) upload_lights(c);
}

bool scene_ready(const orf_context_t* c)
{
    return !c->dirty_geo && !c->dirty_mat && !c->dirty_env && !c->dirty_light && c->d_bvh4.p != nullptr;
}

SceneGPU scene_view(orf_context_t* c)
{
    SceneGPU s;
    std::memset(&s, 0, sizeof(s));
    s.bvh4 = c->d_bvh4.as<float4>();
    s.nodes4 = c->nodes4;
    s.tri = c->d_tri.as<float4>();
    s.shade = c->d_shade.as<int4>();
    s.normals = c->d_normals.as<float4>();
    s.materials = c->d_mats.as<MaterialGPU>();
    s.lights = c->d_lights.as<LightGPU>();
    s.nlights = c->nlights_dev;
    s.tri_lights = c->d_trilights.as<TriLightGPU>();
    s.analytic = c->d_analytic.as<float4>();
    s.nanalytic = (int)(c->analytic.size() / 2);
    s.analytic_prim = c->analytic_prim;
    s.has_env = c->env_w > 0;
    s.env.tex = c->d_env.as<float4>();
    s.env.w = c->env_w;
    s.env.h = c->env_h;
    s.env.marg_pdf = c->d_mpdf.as<float>();
    s.env.marg_cdf = c->d_mcdf.as<float>();
    s.env.cond_pdf = c->d_cpdf.as<float>();
    s.env.cond_cdf = c->d_ccdf.as<float>();
    s.env.sampled = c->env_w > 0 && c->env_is;
    s.max_depth = c->max_depth;
    s.scene_epsilon = c->scene_epsilon;
    s.features = 0u;
    for (const MaterialGPU& m : c->mats) {
        s.features |= material_features(m.kind);
        // the conductor's microsurface walk is compiled with the dielectric code
        if (m.kind == MAT_CONDUCTOR && m.microfacet_model == MULTISCATTERING_MODEL) s.features |= kFeatDielectric;
    }
    // disk / sphere emitters are traced only by the all-feature megakernel variant
    if (!c->analytic.empty()) s.features = kFeatAll;
    s.sss = c->d_sss.as<ScatteringGPU>();
    s.probes.sss = -1;
    if (c->probe_count > 0) {
        s.probes.pos = c->d_probe_pos.as<float4>();
        s.probes.nrm = c->d_probe_nrm.as<float4>();
        s.probes.L = c->d_probe_L.as<float4>();
        s.probes.w12 = c->d_probe_w12.as<float4>();
        s.probes.sn = c->d_probe_sn.as<float4>();
        s.probes.cells = c->d_probe_cells.as<uint2>();
        s.probes.table_mask = c->probe_mask;
        s.probes.cell = c->probe_cell;
        s.probes.count = c->probe_count;
        s.probes.samples = c->probe_samples;
        s.probes.sss = c->probe_sss;
    }
    return s;
}

// Probe pass of the translucent object for sample index `frame` (no-op without probes).
void run_probe_pass(orf_context_t* c, unsigned frame, hipStream_t stream, ScatteringGPU* table)
{
    if (c->probe_count <= 0) return;
    SceneGPU s = scene_view(c);
    if (table) s.sss = table;
    int sss = s.probes.sss;
    s.probes.sss = -1;  // no gathers while the probes are being written
    launch_probe_pass(s, s.sss, sss, frame, stream);
}

void film_clear(orf_context_t* c)
{
    if (c->d_film.p) ORF_HIP_CHECK(hipMemset(c->d_film.p, 0, c->d_film.bytes));
    if (c->d_depth.p) ORF_HIP_CHECK(hipMemset(c->d_depth.p, 0, c->d_depth.bytes));
    if (c->d_counters.p) ORF_HIP_CHECK(hipMemset(c->d_counters.p, 0, c->d_counters.bytes));
}

void film_ensure(orf_context_t* c, int w, int h)
{
    if (w == c->film_w && h == c->film_h && c->d_film.p) return;
    c->d_film.release();
    c->d_film.ensure((size_t)w * h * 4 * sizeof(double));
    c->d_depth.release();
    c->d_depth.ensure((size_t)w * h * sizeof(double));
    c->d_counters.ensure(16 * sizeof(unsigned long long));  // [0..3] counters, [8..15] queue counters
    c->film_w = w;
    c->film_h = h;
    film_clear(c);
}

}  // namespace orf

using orf::ctx_check;

extern "C" orf_status orf_context_create(int device, orf_context* out)
{
    ORF_API_BEGIN
    ORF_REQUIRE(out != nullptr, "orf_context_create: null output");
    int n = 0;
    ORF_HIP_CHECK(hipGetDeviceCount(&n));
    ORF_REQUIRE(device >= 0 && device < n, "orf_context_create: no such device");
    ORF_HIP_CHECK(hipSetDevice(device));
    auto c = std::make_unique<orf_context_t>();
    c->device = device;
    *out = c.release();
    ORF_API_END
}

extern "C" orf_status orf_context_destroy(orf_context ctx)
{
    ORF_API_BEGIN
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        delete ctx;
    }
    ORF_API_END
}

extern "C" orf_status orf_variable_set_float(orf_context ctx, int scope, int handle, const char* name,
                                             const float* v, int n)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(name && v && n > 0 && n <= 4, "orf_variable_set_float: bad arguments");
    if (scope == ORF_SCOPE_CONTEXT) orf::set_context_float(c, name, v, n);
    else if (scope == ORF_SCOPE_MATERIAL) orf::set_material_float(c, handle, name, v, n);
    else if (orf::instance_scope(scope)) orf::set_material_float(c, orf::instance_material(c, handle), name, v, n);
    else ORF_THROW(ORF_ERROR_INVALID_VALUE, "bad variable scope");
    ORF_API_END
}

extern "C" orf_status orf_variable_set_int(orf_context ctx, int scope, int handle, const char* name,
                                           const int32_t* v, int n)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(name && v && n > 0 && n <= 4, "orf_variable_set_int: bad arguments");
    int64_t w[4];
    for (int k = 0; k < n; ++k) w[k] = v[k];
    if (scope == ORF_SCOPE_CONTEXT) orf::set_context_int(c, name, w, n);
    else if (scope == ORF_SCOPE_MATERIAL) orf::set_material_int(c, handle, name, w, n);
    else if (orf::instance_scope(scope)) orf::set_material_int(c, orf::instance_material(c, handle), name, w, n);
    else ORF_THROW(ORF_ERROR_INVALID_VALUE, "bad variable scope");
    ORF_API_END
}

extern "C" orf_status orf_variable_set_uint(orf_context ctx, int scope, int handle, const char* name,
                                            const uint32_t* v, int n)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(name && v && n > 0 && n <= 4, "orf_variable_set_uint: bad arguments");
    int64_t w[4];
    for (int k = 0; k < n; ++k) w[k] = v[k];
    if (scope == ORF_SCOPE_CONTEXT) orf::set_context_int(c, name, w, n);
    else if (scope == ORF_SCOPE_MATERIAL) orf::set_material_int(c, handle, name, w, n);
    else if (orf::instance_scope(scope)) orf::set_material_int(c, orf::instance_material(c, handle), name, w, n);
    else ORF_THROW(ORF_ERROR_INVALID_VALUE, "bad variable scope");
    ORF_API_END
}

extern "C" orf_status orf_variable_set_user_data(orf_context ctx, int scope, int handle, const char* name,
                                                 const void* data, size_t size)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(name && data, "orf_variable_set_user_data: bad arguments");
    ORF_REQUIRE(scope == ORF_SCOPE_MATERIAL || orf::instance_scope(scope),
                "user data is defined on materials and geometry instances");
    if (orf::instance_scope(scope)) handle = orf::instance_material(c, handle);
    std::string nm = name;
    orf::MaterialGPU& m = orf::material_ref(c, handle);
    if (nm == "scattering_properties") {
        // ScatteringMaterialProperties (ORF/structs.h:116-144), read by the programs as given
        ORF_REQUIRE(size == sizeof(orf_scattering_material_properties),
                    "scattering_properties user data must be ScatteringMaterialProperties (260 bytes)");
        ORF_REQUIRE(orf::is_translucent(m.kind), "scattering_properties: not a translucent material");
        orf::set_user_scattering(c->mat_sss[handle], *(const orf_scattering_material_properties*)data);
    } else if (nm == "ior") {
        // MyComplex3: three {double real, double im} (ORF/MyComplex.h:7-16)
        ORF_REQUIRE(size == 6 * sizeof(double), "ior user data must be MyComplex3 (48 bytes)");
        const double* d = (const double*)data;
        double re[3] = {d[0], d[2], d[4]}, im[3] = {d[1], d[3], d[5]};
        orf::conductor_eta_host(re, im, m.eta);
    } else {
        ORF_THROW(ORF_ERROR_INVALID_VALUE, "unknown user-data material variable '" + nm + "'");
    }
    ORF_API_END
}

extern "C" orf_status orf_material_create(orf_context ctx, const char* program, int* material)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(program && material, "orf_material_create: bad arguments");
    int kind = orf::program_kind(program);
    if (kind < 0) ORF_THROW(ORF_ERROR_NOT_FOUND, std::string("no radiance program for '") + program + "'");
    c->mats.push_back(orf::default_material(kind));
    c->mat_sss.push_back(orf::SssParamsHost());
    if (orf::is_translucent(kind) || kind == orf::MAT_TRANSPARENT) c->mats.back().ior = 1.3f;
    if (kind == orf::MAT_ROUGH_DIFFUSE) {  // RoughDiffuseMaterial (ORF/RoughDiffuseMaterial.cpp:7-17)
        orf::MaterialGPU& m = c->mats.back();
        m.diffuse_color = orf::mk3(1.0f);
        m.ior = 1.3f;
        m.normal_distribution = orf::BECKMANN_DISTRIBUTION;
        m.microfacet_model = orf::WALTER_MODEL;
    }
    if (kind == orf::MAT_FLAT) {  // FlatMaterial::FlatMaterial (ORF/FlatMaterial.cpp:5-21)
        orf::MaterialGPU& m = c->mats.back();
        m.diffuse_color = orf::mk3(1.0f);
        m.radiance = orf::mk3(1.0f);
        m.roughness = orf::mk2(0.999f, 0.999f);
        m.normal_distribution = orf::BECKMANN_DISTRIBUTION;
        m.ior = 1.3f;
    }
    if (kind == orf::MAT_ROUGH_TRANSLUCENT) {  // ORF/RoughTranslucentMaterial.cpp:14-18
        c->mats.back().normal_distribution = orf::BECKMANN_DISTRIBUTION;
        c->mats.back().microfacet_model = orf::VISIBLE_NORMALS_MODEL;
    }
    c->dirty_mat = true;
    *material = (int)c->mats.size() - 1;
    ORF_API_END
}

extern "C" orf_status orf_mesh_create(orf_context ctx, const float* verts, int64_t nverts, const int32_t* vindex,
                                      int64_t ntris, const float* normals, int64_t nnormals, const int32_t* nindex,
                                      int material, int* mesh)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(mesh && nverts >= 0 && ntris >= 0, "orf_mesh_create: bad arguments");
    ORF_REQUIRE((nverts == 0 || verts) && (ntris == 0 || vindex), "orf_mesh_create: null arrays");
    ORF_REQUIRE(material >= 0 && material < (int)c->mats.size(), "orf_mesh_create: bad material handle");
    orf::MeshRec m;
    m.v.assign(verts, verts + 3 * nverts);
    m.vi.assign(vindex, vindex + 3 * ntris);
    for (int32_t i : m.vi) ORF_REQUIRE(i >= 0 && i < nverts, "orf_mesh_create: vertex index out of range");
    if (normals && nnormals > 0) {
        m.n.assign(normals, normals + 3 * nnormals);
        if (nindex) m.ni.assign(nindex, nindex + 3 * ntris);
        else {
            ORF_REQUIRE(nnormals == nverts, "orf_mesh_create: normals without nindex need one per vertex");
            m.ni = m.vi;
        }
        for (int32_t i : m.ni) ORF_REQUIRE(i >= 0 && i < nnormals, "orf_mesh_create: normal index out of range");
    }
    m.material = material;
    c->meshes.push_back(std::move(m));
    c->dirty_geo = true;
    *mesh = (int)c->meshes.size() - 1;
    ORF_API_END
}

extern "C" orf_status orf_mesh_set_transform(orf_context ctx, int mesh, const float m[16])
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(mesh >= 0 && mesh < (int)c->meshes.size() && m, "orf_mesh_set_transform: bad arguments");
    std::memcpy(c->meshes[mesh].xf, m, 16 * sizeof(float));
    c->meshes[mesh].has_xf = true;
    c->dirty_geo = true;
    ORF_API_END
}

extern "C" orf_status orf_light_add_point(orf_context ctx, const float position[3], const float radiance[3])
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(position && radiance, "orf_light_add_point: null arguments");
    orf::LightGPU l;
    std::memset(&l, 0, sizeof(l));
    l.type = orf::POINT_LIGHT;
    l.radiance = orf::mk3(radiance[0], radiance[1], radiance[2]);
    l.params = orf::mk3(position[0], position[1], position[2]);
    c->lights.push_back(l);
    c->dirty_light = true;
    ORF_API_END
}

extern "C" orf_status orf_light_add_directional(orf_context ctx, const float direction[3], const float radiance[3])
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(direction && radiance, "orf_light_add_directional: null arguments");
    orf::LightGPU l;
    std::memset(&l, 0, sizeof(l));
    l.type = orf::DIRECTIONAL_LIGHT;
    l.radiance = orf::mk3(radiance[0], radiance[1], radiance[2]);
    // the loader stores the normalised direction (ORF/OptixSceneLoader.cpp:366-372)
    l.params = orf::normalize(orf::mk3(direction[0], direction[1], direction[2]));
    c->lights.push_back(l);
    c->dirty_light = true;
    ORF_API_END
}

extern "C" orf_status orf_light_add_triangles(orf_context ctx, const float* tv, const float* tn, int64_t ntris,
                                              const float radiance[3])
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(tv && radiance && ntris > 0, "orf_light_add_triangles: bad arguments");
    orf::LightGPU l;
    std::memset(&l, 0, sizeof(l));
    l.type = orf::TRIANGLES_AREA_LIGHT;
    l.radiance = orf::mk3(radiance[0], radiance[1], radiance[2]);
    float start = (float)c->tri_lights.size();
    l.params = orf::mk3((float)ntris, start, start + (float)ntris - 1.0f);
    for (int64_t t = 0; t < ntris; ++t) {
        orf::TriLightGPU tl;
        std::memset(&tl, 0, sizeof(tl));
        const float* p = tv + 9 * t;
        tl.v0 = orf::mk3(p[0], p[1], p[2]);
        tl.v1 = orf::mk3(p[3], p[4], p[5]);
        tl.v2 = orf::mk3(p[6], p[7], p[8]);
        if (tn) {
            const float* q = tn + 9 * t;
            tl.n0 = orf::mk3(q[0], q[1], q[2]);
            tl.n1 = orf::mk3(q[3], q[4], q[5]);
            tl.n2 = orf::mk3(q[6], q[7], q[8]);
            tl.has_normals = 1;
        }
        tl.area = (float)(0.5 * (double)orf::length(orf::cross(tl.v1 - tl.v2, tl.v0 - tl.v2)));
        tl.emission = l.radiance;
        c->tri_lights.push_back(tl);
    }
    c->lights.push_back(l);
    c->dirty_light = true;
    ORF_API_END
}

namespace orf {

// Emitter material of a disk or sphere light (the arealight material each
// DiskAreaLight / SphericalLight instance carries, ORF/Light.cpp:476-502, 583-607).
int emitter_material(orf_context_t* c, const float radiance[3])
{
    c->mats.push_back(orf::default_material(orf::MAT_EMITTER));
    c->mats.back().radiance = orf::mk3(radiance[0], radiance[1], radiance[2]);
    c->mat_sss.push_back(orf::SssParamsHost());
    c->dirty_mat = true;
    return (int)c->mats.size() - 1;
}

// A disk or sphere light: the light record, its analytic emitter and the
// emitter's material.
void add_analytic_light(orf_context_t* c, unsigned type, const float position[3], float radius, orf::f3 normal,
                               const float radiance[3], int* material)
{
    const int m = emitter_material(c, radiance);
    const int idx = (int)(c->analytic.size() / 2);
    c->analytic.push_back(make_float4(position[0], position[1], position[2], radius));
    c->analytic.push_back(make_float4(normal.x, normal.y, normal.z, type == orf::DISK_LIGHT ? 0.0f : 1.0f));
    c->analytic_mat.push_back(m);
    orf::LightGPU l;
    std::memset(&l, 0, sizeof(l));
    l.type = type;
    l.radiance = orf::mk3(radiance[0], radiance[1], radiance[2]);
    l.params = orf::mk3(position[0], position[1], position[2]);
    l.pad = orf::mk3(radius, (float)idx, 0.0f);
    c->lights.push_back(l);
    c->dirty_light = true;
    c->dirty_geo = true;
    if (material) *material = m;
}

}  // namespace orf

extern "C" orf_status orf_light_add_disk(orf_context ctx, const float position[3], float radius, float theta,
                                         float phi, const float radiance[3], int* emitter_material)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(position && radiance, "orf_light_add_disk: null arguments");
    ORF_REQUIRE(radius > 0.0f && std::isfinite(radius), "orf_light_add_disk: radius must be positive and finite");
    // DiskAreaLight::computeNormal (ORF/Light.cpp:469-474): theta from +y, phi about y, in degrees
    const float t = theta * orf::kPi / 180.0f, p = phi * orf::kPi / 180.0f;
    const orf::f3 n = orf::mk3(orf_sinf(t) * orf_sinf(p), orf_cosf(t), orf_sinf(t) * orf_cosf(p));
    orf::add_analytic_light(c, orf::DISK_LIGHT, position, radius, n, radiance, emitter_material);
    ORF_API_END
}

extern "C" orf_status orf_light_add_sphere(orf_context ctx, const float position[3], float radius,
                                           const float radiance[3], int* emitter_material)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(position && radiance, "orf_light_add_sphere: null arguments");
    ORF_REQUIRE(radius > 0.0f && std::isfinite(radius), "orf_light_add_sphere: radius must be positive and finite");
    orf::add_analytic_light(c, orf::SPHERICAL_LIGHT, position, radius, orf::mk3(0.0f, 0.0f, 0.0f), radiance,
                       emitter_material);
    ORF_API_END
}

extern "C" orf_status orf_lights_clear(orf_context ctx)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    c->lights.clear();
    c->tri_lights.clear();
    if (!c->analytic.empty()) c->dirty_geo = true;  // the emitters go with their lights
    c->analytic.clear();
    c->analytic_mat.clear();
    c->dirty_light = true;
    ORF_API_END
}

extern "C" orf_status orf_envmap_set(orf_context ctx, const float* rgba, int width, int height, int importance_sample)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(rgba && width > 0 && height > 0, "orf_envmap_set: bad arguments");
    c->env_tex.resize((size_t)width * height);
    std::memcpy(c->env_tex.data(), rgba, c->env_tex.size() * sizeof(float4));
    c->env_w = width;
    c->env_h = height;
    c->env_is = importance_sample != 0;
    c->dirty_env = true;
    ORF_API_END
}

extern "C" orf_status orf_envmap_tables(orf_context ctx, float* mpdf, float* mcdf, float* cpdf, float* ccdf)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(c->env_w > 0 && c->env_is, "orf_envmap_tables: no importance-sampled environment");
    orf::prepare(c, nullptr);
    size_t w = c->env_w, h = c->env_h;
    if (mpdf) ORF_HIP_CHECK(hipMemcpy(mpdf, c->d_mpdf.p, h * 4, hipMemcpyDeviceToHost));
    if (mcdf) ORF_HIP_CHECK(hipMemcpy(mcdf, c->d_mcdf.p, h * 4, hipMemcpyDeviceToHost));
    if (cpdf) ORF_HIP_CHECK(hipMemcpy(cpdf, c->d_cpdf.p, w * h * 4, hipMemcpyDeviceToHost));
    if (ccdf) ORF_HIP_CHECK(hipMemcpy(ccdf, c->d_ccdf.p, w * h * 4, hipMemcpyDeviceToHost));
    ORF_API_END
}

extern "C" orf_status orf_accel_mark_dirty(orf_context ctx)
{
    ORF_API_BEGIN
    ctx_check(ctx)->dirty_geo = true;
    ORF_API_END
}

extern "C" orf_status orf_accel_build(orf_context ctx, int64_t* stats)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    orf::prepare(c, nullptr);
    if (stats) {
        stats[0] = c->bvh_stats.nodes;
        stats[1] = c->bvh_stats.leaves;
        stats[2] = c->bvh_stats.max_depth;
        stats[3] = c->bvh_stats.max_leaf;
        stats[4] = c->bvh_stats.tris;
        stats[5] = c->bvh_stats.degenerate;
    }
    ORF_API_END
}

extern "C" orf_status orf_launch(orf_context ctx, int width, int height, uint32_t frame0, uint32_t spp, void* stream)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(width > 0 && height > 0 && width <= 16384 && height <= 16384, "orf_launch: bad film size");
    ORF_REQUIRE(c->max_depth <= 4096, "orf_launch: max_depth too large");
    hipStream_t st = (hipStream_t)stream;
    orf::prepare(c, st);
    orf::film_ensure(c, width, height);
    orf::CameraGPU cam = c->cam;
    if (!c->patch_set) {
        cam.patch_origin[0] = cam.patch_origin[1] = 0;
        cam.patch_dims[0] = (unsigned)width;
        cam.patch_dims[1] = (unsigned)height;
    }
    orf::FilmGPU film{c->d_film.as<double>(), c->d_counters.as<unsigned long long>(), width, height,
                      c->d_counters.as<unsigned long long>() + 8};
    if (c->integrator == 1) {  // DepthTracer (ORF/Integrator.cpp:100-118)
        orf::launch_depth(orf::scene_view(c), cam, film, c->d_depth.as<double>(), frame0, spp, st);
        return ORF_SUCCESS;
    }
    orf::run_probe_pass(c, frame0, st);
    orf::launch_megakernel(orf::scene_view(c), cam, film, nullptr, frame0, spp, st);
    ORF_API_END
}

extern "C" orf_status orf_integrator_set(orf_context ctx, int integrator)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(integrator == ORF_INTEGRATOR_PATH_TRACER || integrator == ORF_INTEGRATOR_DEPTH,
                "orf_integrator_set: unknown integrator");
    c->integrator = integrator;
    ORF_API_END
}

extern "C" orf_status orf_film_clear(orf_context ctx)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_HIP_CHECK(hipDeviceSynchronize());
    orf::film_clear(c);
    ORF_API_END
}

extern "C" orf_status orf_film_read_sums(orf_context ctx, double* sums)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(sums && c->d_film.p, "orf_film_read_sums: no film");
    ORF_HIP_CHECK(hipDeviceSynchronize());
    ORF_HIP_CHECK(hipMemcpy(sums, c->d_film.p, (size_t)c->film_w * c->film_h * 4 * sizeof(double),
                            hipMemcpyDeviceToHost));
    ORF_API_END
}

extern "C" orf_status orf_film_display(orf_context ctx, float* rgba32_device, void* stream)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(c->d_film.p, "orf_film_display: no film");
    ORF_REQUIRE(rgba32_device, "orf_film_display: null buffer");
    const long long n = (long long)c->film_w * c->film_h;
    const double* depth = c->integrator == 1 ? c->d_depth.as<double>() : nullptr;
    ORF_REQUIRE(c->integrator != 1 || depth, "orf_film_display: no depth film");
    orf::launch_film_display(c->d_film.as<double>(), depth, n, (float4*)rgba32_device, (hipStream_t)stream);
    ORF_API_END
}

extern "C" orf_status orf_film_read(orf_context ctx, double* rgba64, float* rgba32)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(c->d_film.p, "orf_film_read: no film");
    size_t n = (size_t)c->film_w * c->film_h;
    std::vector<double> s(4 * n);
    ORF_HIP_CHECK(hipDeviceSynchronize());
    ORF_HIP_CHECK(hipMemcpy(s.data(), c->d_film.p, s.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<double> dep;
    if (c->integrator == 1) {
        dep.resize(n);
        ORF_HIP_CHECK(hipMemcpy(dep.data(), c->d_depth.p, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (size_t i = 0; i < n; ++i) {
        double cnt = s[4 * i + 3];
        for (int k = 0; k < 3; ++k) {
            double v = cnt > 0 ? s[4 * i + k] / cnt : 0.0;
            if (rgba64) rgba64[4 * i + k] = v;
            if (rgba32) rgba32[4 * i + k] = (float)v;
        }
        // alpha: 1 for the path tracer (CU/path_tracer.cu:67), the mean hit
        // distance for the depth tracer (CU/depth_tracer.cu:52)
        double a = dep.empty() ? 1.0 : (cnt > 0 ? dep[i] / cnt : 0.0);
        if (rgba64) rgba64[4 * i + 3] = a;
        if (rgba32) rgba32[4 * i + 3] = (float)a;
    }
    ORF_API_END
}

extern "C" orf_status orf_counters(orf_context ctx, uint64_t out[4])
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(out, "orf_counters: null output");
    if (!c->d_counters.p) {
        for (int k = 0; k < 4; ++k) out[k] = 0;
        return ORF_SUCCESS;
    }
    ORF_HIP_CHECK(hipDeviceSynchronize());
    ORF_HIP_CHECK(hipMemcpy(out, c->d_counters.p, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    ORF_API_END
}

extern "C" orf_status orf_scattering_properties(orf_context ctx, int material, float* out, int n)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(out && n >= 64, "orf_scattering_properties: need 64 floats");
    ORF_REQUIRE(material >= 0 && material < (int)c->mats.size(), "orf_scattering_properties: bad material");
    orf::ScatteringGPU p = orf::scattering_coefficients(c->mat_sss[material]);
    const orf::f3* v[] = {&p.absorption, &p.scattering, &p.meancosine, &p.extinction, &p.reducedExtinction,
                          &p.deltaEddExtinction, &p.D, &p.transport, &p.C_phi, &p.C_phi_inv, &p.C_E, &p.A,
                          &p.albedo, &p.reducedAlbedo, &p.de, &p.three_D, &p.two_A_de, &p.global_coeff,
                          &p.one_over_three_ext};
    int k = 0;
    out[k++] = p.relative_ior;
    out[k++] = (float)p.dipole_model;
    for (const orf::f3* x : v) {
        out[k++] = x->x;
        out[k++] = x->y;
        out[k++] = x->z;
    }
    out[k++] = p.mean_transport;
    while (k < n) out[k++] = 0.0f;
    ORF_API_END
}

extern "C" orf_status orf_probes_read(orf_context ctx, int64_t* count, double* area, float* cell, float* rmin,
                                      float* pos, float* nrm, float* L, float* w12, float* sn)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    orf::prepare(c, nullptr);
    if (count) *count = c->probe_count;
    if (area) *area = c->probe_area;
    if (cell) *cell = c->probe_cell;
    if (rmin) *rmin = c->probe_rmin;
    size_t bytes = (size_t)c->probe_count * sizeof(float4);
    ORF_HIP_CHECK(hipDeviceSynchronize());
    if (c->probe_count > 0) {
        if (pos) ORF_HIP_CHECK(hipMemcpy(pos, c->d_probe_pos.p, bytes, hipMemcpyDeviceToHost));
        if (nrm) ORF_HIP_CHECK(hipMemcpy(nrm, c->d_probe_nrm.p, bytes, hipMemcpyDeviceToHost));
        if (L) ORF_HIP_CHECK(hipMemcpy(L, c->d_probe_L.p, bytes, hipMemcpyDeviceToHost));
        if (w12) ORF_HIP_CHECK(hipMemcpy(w12, c->d_probe_w12.p, bytes, hipMemcpyDeviceToHost));
        if (sn) ORF_HIP_CHECK(hipMemcpy(sn, c->d_probe_sn.p, bytes, hipMemcpyDeviceToHost));
    }
    ORF_API_END
}

extern "C" orf_status orf_probe_pass(orf_context ctx, uint32_t frame, void* stream)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    hipStream_t st = (hipStream_t)stream;
    orf::prepare(c, st);
    orf::run_probe_pass(c, frame, st);
    ORF_HIP_CHECK(hipStreamSynchronize(st));
    ORF_API_END
}
