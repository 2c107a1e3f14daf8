// orf_context.h -- the render context behind the C-ABI handle (internal).
#pragma once

#include <string>
#include <vector>

#include "orf_bvh.h"
#include "orf_devbuf.h"
#include "orf_probes_host.h"
#include "orf_render.h"
#include "orf_wavefront_host.h"

namespace orf {

struct MeshRec {
    std::vector<float> v, n;
    std::vector<int32_t> vi, ni;
    int material = 0;
    float xf[16];
    bool has_xf = false;
    bool own_material = false;  // `material` is this mesh's private copy (instance-scope variables)
};

}  // namespace orf

struct orf_context_t {
    int device = 0;
    // context variables (defaults of ORF/Integrator.cpp:6-13)
    int max_depth = 50;
    float scene_epsilon = 1e-4f;
    orf::CameraGPU cam{};
    bool patch_set = false;
    unsigned frame = 0;
    float exception_color[3] = {1.0f, 0.0f, 0.0f};
    // scene records
    std::vector<orf::MaterialGPU> mats;
    std::vector<orf::SssParamsHost> mat_sss;  // translucent parameters, per material handle
    std::vector<orf::MeshRec> meshes;
    std::vector<orf::LightGPU> lights;
    std::vector<orf::TriLightGPU> tri_lights;
    std::vector<float4> analytic;        // disk / sphere emitters: 2 float4 each
    std::vector<int> analytic_mat;       // their emitter material handles
    int analytic_prim = 0;               // shade index of the first emitter
    std::vector<float4> env_tex;
    int env_w = 0, env_h = 0;
    bool env_is = false;
    bool dirty_geo = true, dirty_mat = true, dirty_light = true, dirty_env = true;
    orf::BvhStats bvh_stats;
    int nodes4 = 0;  // 4-wide BVH nodes on the device
    // device copies
    orf::DevBuf d_bvh4, d_tri, d_shade, d_normals, d_mats, d_lights, d_trilights, d_analytic;
    orf::DevBuf d_env, d_mpdf, d_mcdf, d_cpdf, d_ccdf, d_envscratch;
    orf::DevBuf d_film, d_counters;
    // translucent material: coefficients and surface probes
    orf::DevBuf d_sss, d_probe_pos, d_probe_nrm, d_probe_L, d_probe_w12, d_probe_sn, d_probe_cells;
    int probe_count_req = 0;       // context variable probe_count (0: no probes)
    int probe_samples = 4;         // context variable probe_samples (gather samples per hit)
    unsigned probe_seed = 11;      // context variable probe_seed
    int probe_sss = -1;            // material handle whose mesh carries the probes
    int probe_count = 0;           // probes generated
    unsigned probe_mask = 0;
    float probe_cell = 0.0f, probe_rmin = 0.0f;
    double probe_area = 0.0;
    int film_w = 0, film_h = 0;
    int nlights_dev = 0;
    // batched renders
    orf::DevBuf d_batch_renders, d_batch_remap, d_batch_film, d_batch_stats, d_batch_items;
    std::vector<orf::WfRender> batch_renders;
    long long batch_pixels = 0;
    orf::WavefrontPool wf_pool;
    double last_batch_seconds = 0.0;
    // differentiable renders
    orf::DevBuf d_diff_views, d_diff_stats, d_diff_sss, d_diff_partial;
    long long last_batch_iterations = 0;
    // ray generation program: 0 path tracer, 1 depth tracer; the depth film
    int integrator = 0;
    orf::DevBuf d_depth;
    // environment bound through orf_variable_set_texture: importance sampled?
    bool envmap_importance = true;
    // textures of orf_texture_create (FP32 RGBA, row 0 first)
    struct Texture {
        std::vector<float4> texels;
        int w = 0, h = 0, wrap = 0, filter = 0;
    };
    std::vector<Texture> textures;
    // the scene document last read by orf_scene_load_json (saved back by orf_scene_save)
    std::string scene_doc;
};

namespace orf {
orf_context_t* ctx_check(orf_context ctx);
void prepare(orf_context_t* c, hipStream_t stream);
// True when every scene change is on the device (prepare() would do nothing).
bool scene_ready(const orf_context_t* c);
SceneGPU scene_view(orf_context_t* c);
// table: the ScatteringGPU table to use (null: the context's)
void run_probe_pass(orf_context_t* c, unsigned frame, hipStream_t stream, ScatteringGPU* table = nullptr);
void film_ensure(orf_context_t* c, int w, int h);
void film_clear(orf_context_t* c);
MaterialGPU default_material(int kind);
int program_kind(const std::string& prog);
MaterialGPU& material_ref(orf_context_t* c, int handle);
void set_material_float(orf_context_t* c, int handle, const std::string& name, const float* v, int n);
void set_material_int(orf_context_t* c, int handle, const std::string& name, const int64_t* v, int n);
int emitter_material(orf_context_t* c, const float radiance[3]);
void add_analytic_light(orf_context_t* c, unsigned type, const float position[3], float radius, f3 normal,
                        const float radiance[3], int* material);
// Geometry and instance scopes address a mesh's single instance.
inline bool instance_scope(int scope) { return scope == ORF_SCOPE_GEOMETRY || scope == ORF_SCOPE_INSTANCE; }
// The mesh's private material (copied from its shared one on first use).
int instance_material(orf_context_t* c, int mesh);
// ScatteringMaterialProperties as given by the caller become the material's coefficients.
void set_user_scattering(SssParamsHost& h, const orf_scattering_material_properties& p);
}  // namespace orf

