// orf_boundary.cpp -- the rest of the optix:: host API surface the reference's
// scene classes use, on the render context: geometry / instance variable
// scopes, Matrix4x4 variables, user-format light buffers, per-ray-type program
// binding and texture samplers.
#include <cmath>
#include <cstring>
#include <string>

#include "orf_common.h"
#include "orf_context.h"
#include "orf_math.h"

namespace orf {

int instance_material(orf_context_t* c, int mesh)
{
    ORF_REQUIRE(mesh >= 0 && mesh < (int)c->meshes.size(), "bad mesh handle for a geometry / instance variable");
    MeshRec& m = c->meshes[mesh];
    if (!m.own_material) {
        // the instance's own variables shadow its material's from here on
        MaterialGPU copy = c->mats[m.material];
        SssParamsHost sss = c->mat_sss[m.material];
        c->mats.push_back(copy);
        c->mat_sss.push_back(sss);
        m.material = (int)c->mats.size() - 1;
        m.own_material = true;
        c->dirty_mat = true;
        c->dirty_geo = true;  // shade records carry the material id
    }
    return m.material;
}

void set_user_scattering(SssParamsHost& h, const orf_scattering_material_properties& p)
{
    ScatteringGPU u;
    std::memset(&u, 0, sizeof(u));
    auto v3 = [](const float* a) { return mk3(a[0], a[1], a[2]); };
    u.relative_ior = p.relative_ior;
    u.absorption = v3(p.absorption);
    u.scattering = v3(p.scattering);
    u.meancosine = v3(p.meancosine);
    u.extinction = v3(p.extinction);
    u.reducedExtinction = v3(p.reducedExtinction);
    u.deltaEddExtinction = v3(p.deltaEddExtinction);
    u.D = v3(p.D);
    u.transport = v3(p.transport);
    u.C_phi = v3(p.C_phi);
    u.C_phi_inv = v3(p.C_phi_inv);
    u.C_E = v3(p.C_E);
    u.A = v3(p.A);
    u.albedo = v3(p.albedo);
    u.reducedAlbedo = v3(p.reducedAlbedo);
    u.de = v3(p.de);
    u.three_D = v3(p.three_D);
    u.two_A_de = v3(p.two_A_de);
    u.global_coeff = v3(p.global_coeff);
    u.one_over_three_ext = v3(p.one_over_three_ext);
    u.mean_transport = p.mean_transport;
    h.user = u;
    h.has_user = true;
}

}  // namespace orf

using orf::ctx_check;

extern "C" orf_status orf_variable_set_matrix4x4(orf_context ctx, int scope, int handle, const char* name,
                                                 int transpose, const float m[16])
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(name && m, "orf_variable_set_matrix4x4: bad arguments");
    float r[16];  // row-major
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r[4 * i + j] = transpose ? m[4 * j + i] : m[4 * i + j];
    for (float x : r) ORF_REQUIRE(std::isfinite(x), "orf_variable_set_matrix4x4: non-finite entry");
    const std::string nm = name;
    if (scope == ORF_SCOPE_CONTEXT) {
        // the probe program's current object (CU/sample_camera.cu:45-46): the probe
        // pass here takes every translucent mesh's own transform
        ORF_REQUIRE(nm == "transform_matrix" || nm == "normal_matrix",
                    "unknown Matrix4x4 context variable '" + nm + "'");
    } else if (orf::instance_scope(scope)) {
        ORF_REQUIRE(handle >= 0 && handle < (int)c->meshes.size(), "orf_variable_set_matrix4x4: bad mesh handle");
        orf::MeshRec& mesh = c->meshes[handle];
        if (nm == "transform_matrix") {
            std::memcpy(mesh.xf, r, sizeof(r));
            mesh.has_xf = true;
            c->dirty_geo = true;
        } else if (nm == "inverse_transform_matrix") {
            float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
            const float* M = mesh.has_xf ? mesh.xf : id;
            double scale = 1.0, err = 0.0;
            for (int k = 0; k < 16; ++k) scale = std::fmax(scale, std::fabs((double)M[k]));
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < 4; ++k) s += (double)M[4 * i + k] * (double)r[4 * k + j];
                    err = std::fmax(err, std::fabs(s - (i == j ? 1.0 : 0.0)));
                }
            ORF_REQUIRE(err <= 1e-4 * scale, "inverse_transform_matrix is not the inverse of the mesh's transform");
        } else {
            ORF_THROW(ORF_ERROR_INVALID_VALUE, "unknown Matrix4x4 instance variable '" + nm + "'");
        }
    } else {
        ORF_THROW(ORF_ERROR_INVALID_VALUE, "Matrix4x4 variables are defined on the context and on instances");
    }
    ORF_API_END
}

extern "C" orf_status orf_buffer_set(orf_context ctx, const char* name, const void* data, size_t element_size,
                                     size_t count)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(name && (count == 0 || data), "orf_buffer_set: bad arguments");
    ORF_REQUIRE(count <= (1u << 24), "orf_buffer_set: too many elements");
    const std::string nm = name;
    if (nm == "light_buffer") {
        ORF_REQUIRE(element_size == sizeof(orf_light_struct), "light_buffer elements are LightStruct (40 bytes)");
        const orf_light_struct* L = (const orf_light_struct*)data;
        // validate first: a refused buffer leaves the light list as it was
        for (size_t i = 0; i < count; ++i) {
            ORF_REQUIRE(L[i].light_type <= orf::SPHERICAL_LIGHT,
                        "light_buffer[" + std::to_string(i) + "]: unknown light_type");
            if (L[i].light_type == orf::DISK_LIGHT || L[i].light_type == orf::SPHERICAL_LIGHT)
                ORF_REQUIRE(L[i].padding[0] > 0.0f && std::isfinite(L[i].padding[0]),
                            "light_buffer[" + std::to_string(i) + "]: radius must be positive and finite");
        }
        c->lights.clear();
        if (!c->analytic.empty()) c->dirty_geo = true;
        c->analytic.clear();
        c->analytic_mat.clear();
        for (size_t i = 0; i < count; ++i) {
            const orf_light_struct& e = L[i];
            if (e.light_type == orf::DISK_LIGHT || e.light_type == orf::SPHERICAL_LIGHT) {
                // DiskLightStruct {radius, phi, theta} / SphericalLightStruct {radius} (ORF/structs.h:76-93),
                // with the emitter the reference's light geometry carries
                orf::f3 n = orf::mk3(0.0f);
                if (e.light_type == orf::DISK_LIGHT) {
                    const float t = e.padding[2] * orf::kPi / 180.0f, p = e.padding[1] * orf::kPi / 180.0f;
                    n = orf::mk3(orf_sinf(t) * orf_sinf(p), orf_cosf(t), orf_sinf(t) * orf_cosf(p));
                }
                orf::add_analytic_light(c, e.light_type, e.light_parameters, e.padding[0], n, e.emitted_radiance,
                                        nullptr);
                continue;
            }
            orf::LightGPU l;
            std::memset(&l, 0, sizeof(l));
            l.type = e.light_type;
            l.radiance = orf::mk3(e.emitted_radiance[0], e.emitted_radiance[1], e.emitted_radiance[2]);
            l.params = orf::mk3(e.light_parameters[0], e.light_parameters[1], e.light_parameters[2]);
            l.pad = orf::mk3(e.padding[0], e.padding[1], e.padding[2]);
            c->lights.push_back(l);
        }
        c->dirty_light = true;
    } else if (nm == "triangle_light_buffer") {
        ORF_REQUIRE(element_size == sizeof(orf_triangle_light),
                    "triangle_light_buffer elements are TriangleLight (92 bytes)");
        static_assert(sizeof(orf_triangle_light) == sizeof(orf::TriLightGPU), "TriangleLight layout");
        c->tri_lights.resize(count);
        if (count) std::memcpy(c->tri_lights.data(), data, count * sizeof(orf_triangle_light));
        c->dirty_light = true;
    } else {
        ORF_THROW(ORF_ERROR_INVALID_VALUE, "unknown buffer '" + nm + "'");
    }
    ORF_API_END
}

extern "C" orf_status orf_material_set_program(orf_context ctx, int material, int ray_type, const char* program_file,
                                               const char* entry)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(program_file && entry, "orf_material_set_program: bad arguments");
    orf::MaterialGPU& m = orf::material_ref(c, material);
    std::string prog = program_file, ent = entry;
    if (prog.size() > 3 && prog.compare(prog.size() - 3, 3, ".cu") == 0) prog = prog.substr(0, prog.size() - 3);
    if (ray_type == ORF_RAY_RADIANCE && ent == "closest_hit") {
        const int kind = orf::program_kind(prog);
        if (kind < 0) ORF_THROW(ORF_ERROR_NOT_FOUND, "no radiance closest hit for '" + prog + "'");
        if (kind != m.kind) {
            m.kind = kind;
            c->dirty_geo = true;  // the shadow pass-through flag follows the radiance program
        }
    } else if (ray_type == ORF_RAY_SHADOW && ent == "any_hit") {
        // every reference any-hit terminates the shadow ray except normal_shader's
        if (orf::program_kind(prog) < 0) ORF_THROW(ORF_ERROR_NOT_FOUND, "no shadow any hit for '" + prog + "'");
        ORF_REQUIRE((prog == "normal_shader") == (m.kind == orf::MAT_NORMAL),
                    "shadow any hit '" + prog + "' does not match the material's radiance program");
    } else if (ray_type == ORF_RAY_DEPTH && ent == "closest_hit") {
        if (prog != "depth_shader" && prog != "rough_depth_shader")
            ORF_THROW(ORF_ERROR_NOT_FOUND, "no depth closest hit for '" + prog + "'");
    } else {
        ORF_THROW(ORF_ERROR_NOT_FOUND, "no program '" + ent + "' for ray type " + std::to_string(ray_type));
    }
    ORF_API_END
}

extern "C" orf_status orf_texture_create(orf_context ctx, const float* rgba, int width, int height, int wrap_mode,
                                         int filter_mode, int* texture)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    ORF_REQUIRE(rgba && texture && width > 0 && height > 0 && width <= 65536 && height <= 65536,
                "orf_texture_create: bad arguments");
    ORF_REQUIRE(wrap_mode == ORF_WRAP_REPEAT || wrap_mode == ORF_WRAP_CLAMP_TO_EDGE, "orf_texture_create: bad wrap mode");
    ORF_REQUIRE(filter_mode == ORF_FILTER_LINEAR || filter_mode == ORF_FILTER_NEAREST,
                "orf_texture_create: bad filter mode");
    orf_context_t::Texture t;
    t.w = width;
    t.h = height;
    t.wrap = wrap_mode;
    t.filter = filter_mode;
    t.texels.resize((size_t)width * height);
    std::memcpy(t.texels.data(), rgba, t.texels.size() * sizeof(float4));
    c->textures.push_back(std::move(t));
    *texture = (int)c->textures.size() - 1;
    ORF_API_END
}

extern "C" orf_status orf_variable_set_texture(orf_context ctx, int scope, int handle, const char* name, int texture)
{
    ORF_API_BEGIN
    orf_context_t* c = ctx_check(ctx);
    (void)handle;
    ORF_REQUIRE(name, "orf_variable_set_texture: bad arguments");
    ORF_REQUIRE(texture >= 0 && texture < (int)c->textures.size(), "orf_variable_set_texture: bad texture handle");
    const std::string nm = name;
    ORF_REQUIRE(scope == ORF_SCOPE_CONTEXT && nm == "envmap", "unknown texture variable '" + nm + "'");
    const orf_context_t::Texture& t = c->textures[texture];
    // the lat-long lookup filters bilinearly with wrap-around (ORF/HDRLoader.cpp:231-286)
    ORF_REQUIRE(t.wrap == ORF_WRAP_REPEAT && t.filter == ORF_FILTER_LINEAR,
                "envmap: the environment texture must use REPEAT wrapping and LINEAR filtering");
    c->env_tex = t.texels;
    c->env_w = t.w;
    c->env_h = t.h;
    c->env_is = c->envmap_importance;
    c->dirty_env = true;
    ORF_API_END
}
