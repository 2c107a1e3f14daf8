// orf_caller.cpp -- a compiled C++ caller of liborf_rt.so that drives a scene
// through the C-ABI the way the reference's scene classes drive OptiX, call for
// call (the reference call site is named above each group), using the
// checker of the integration guide: every status is rethrown as an exception.
//
// usage: orf_caller <scene.bin> <film.bin>
//   scene.bin (little-endian): int32 W, H, spp, n_ground_v, n_ground_t, n_sphere_v,
//   n_sphere_t, env_w, env_h; float32 eye[3] U[3] V[3] W[3]; ground verts, int32
//   ground tris; sphere verts, sphere normals, int32 sphere tris; 4 light-quad
//   verts, int32 2 light tris; env RGBA.
//   film.bin: W*H*4 doubles (sum R, G, B, count per pixel, row 0 = bottom).
// The scene: a Lambert ground, a GGX conductor sphere moved by an instance
// Matrix4x4, a whole-milk translucent sphere (ScatteringMaterialProperties set
// as user data, computed here as RoughTranslucentMaterial::computeCoefficients
// does) moved by a transposed Matrix4x4, a quad emitter whose radiance is an
// instance variable, a point light and the quad's triangles through the raw
// light buffers, and an environment texture sampler.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "orf_rt.h"

struct OrfException : std::runtime_error {  // role of optix::Exception
    orf_status code;
    explicit OrfException(orf_status s) : std::runtime_error(std::string("orf_rt: ") + orf_last_error()), code(s) {}
};
static void orf_check(orf_status s)
{
    if (s != ORF_SUCCESS) throw OrfException(s);
}

// Fresnel-moment fits and C_phi / C_E (ORF/Fresnel.h:93-124)
static float two_C1(float n)
{
    if (n >= 1.0f) return -9.23372f + n * (22.2272f + n * (-20.9292f + n * (10.2291f + n * (-2.54396f + n * 0.254913f))));
    return 0.919317f + n * (-3.4793f + n * (6.75335f + n * (-7.80989f + n * (4.98554f - n * 1.36881f))));
}
static float three_C2(float n)
{
    if (n >= 1.0f) {
        float r = -1641.1f + n * (1213.67f + n * (-568.556f + n * (164.798f + n * (-27.0181f + n * 1.91826f))));
        return r + (((135.926f / n) - 656.175f) / n + 1376.53f) / n;
    }
    return 0.828421f + n * (-2.62051f + n * (3.36231f + n * (-1.95284f + n * (0.236494f + n * 0.145787f))));
}
static float C_phi(float n) { return 0.25f * (1.0f - two_C1(n)); }
static float C_E(float n) { return 0.5f * (1.0f - three_C2(n)); }

// RoughTranslucentMaterial::computeCoefficients (ORF/RoughTranslucentMaterial.cpp:366-395), outside ior 1
static orf_scattering_material_properties milk_properties(const float sa[3], const float ss[3], const float g[3],
                                                          float scale, float ior)
{
    orf_scattering_material_properties p;
    std::memset(&p, 0, sizeof(p));
    p.relative_ior = ior;
    const float pi = 3.14159265358979323846f;
    for (int c = 0; c < 3; ++c) {
        p.ior_real[c] = ior;
        p.absorption[c] = sa[c] * scale;
        p.scattering[c] = ss[c] * scale;
        p.meancosine[c] = g[c];
        const float red = p.scattering[c] * (1.0f - g[c]);
        p.extinction[c] = p.scattering[c] + p.absorption[c];
        p.reducedExtinction[c] = red + p.absorption[c];
        p.deltaEddExtinction[c] = p.scattering[c] * (1.0f - g[c] * g[c]) + p.absorption[c];
        p.D[c] = 1.0f / (3.0f * p.reducedExtinction[c]);
        p.transport[c] = std::sqrt(p.absorption[c] / p.D[c]);
        p.C_phi[c] = C_phi(ior);
        p.C_phi_inv[c] = C_phi(1.0f / ior);
        p.C_E[c] = C_E(ior);
        p.albedo[c] = p.scattering[c] / p.extinction[c];
        p.reducedAlbedo[c] = red / p.reducedExtinction[c];
        p.de[c] = 2.131f * p.D[c] / std::sqrt(p.reducedAlbedo[c]);
        p.A[c] = (1.0f - p.C_E[c]) / (2.0f * p.C_phi[c]);
        p.three_D[c] = 3.0f * p.D[c];
        p.two_A_de[c] = 2.0f * p.A[c] * p.de[c];
        p.global_coeff[c] = 1.0f / (4.0f * p.C_phi_inv[c]) * 1.0f / (4.0f * pi * pi);
        p.one_over_three_ext[c] = 1.0f / (3.0f * p.extinction[c]);
    }
    p.mean_transport = (p.transport[0] + p.transport[1] + p.transport[2]) / 3.0f;
    return p;
}

template <class T>
static std::vector<T> take(FILE* f, size_t n)
{
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short scene file");
    return v;
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s scene.bin film.bin\n", argv[0]);
        return 2;
    }
    try {
        FILE* f = std::fopen(argv[1], "rb");
        if (!f) throw std::runtime_error("cannot open scene file");
        auto hdr = take<int32_t>(f, 9);
        const int W = hdr[0], H = hdr[1], spp = hdr[2], ngv = hdr[3], ngt = hdr[4], nsv = hdr[5], nst = hdr[6];
        const int ew = hdr[7], eh = hdr[8];
        auto cam = take<float>(f, 12);
        auto gv = take<float>(f, 3 * (size_t)ngv);
        auto gt = take<int32_t>(f, 3 * (size_t)ngt);
        auto sv = take<float>(f, 3 * (size_t)nsv);
        auto sn = take<float>(f, 3 * (size_t)nsv);
        auto st = take<int32_t>(f, 3 * (size_t)nst);
        auto lv = take<float>(f, 12);
        auto lt = take<int32_t>(f, 6);
        auto env = take<float>(f, 4 * (size_t)ew * eh);
        std::fclose(f);

        // ORF/OptixScene.cpp:96-117  Context::create, ray types, entry points
        orf_context ctx;
        orf_check(orf_context_create(0, &ctx));
        // ORF/Integrator.cpp:6-24 and ORF/Camera.cpp:83-103
        int32_t depth = 8;
        orf_check(orf_variable_set_int(ctx, ORF_SCOPE_CONTEXT, 0, "max_depth", &depth, 1));
        float eps = 1e-4f;
        orf_check(orf_variable_set_float(ctx, ORF_SCOPE_CONTEXT, 0, "scene_epsilon", &eps, 1));
        orf_check(orf_variable_set_float(ctx, ORF_SCOPE_CONTEXT, 0, "eye", &cam[0], 3));
        orf_check(orf_variable_set_float(ctx, ORF_SCOPE_CONTEXT, 0, "U", &cam[3], 3));
        orf_check(orf_variable_set_float(ctx, ORF_SCOPE_CONTEXT, 0, "V", &cam[6], 3));
        orf_check(orf_variable_set_float(ctx, ORF_SCOPE_CONTEXT, 0, "W", &cam[9], 3));

        // DiffuseMaterial::initPrograms: closest hit (radiance), any hit (shadow), depth
        int ground;
        orf_check(orf_material_create(ctx, "diffuse_shader.cu", &ground));
        orf_check(orf_material_set_program(ctx, ground, ORF_RAY_SHADOW, "diffuse_shader.cu", "any_hit"));
        orf_check(orf_material_set_program(ctx, ground, ORF_RAY_DEPTH, "depth_shader.cu", "closest_hit"));
        float half[3] = {0.5f, 0.5f, 0.5f};
        orf_check(orf_variable_set_float(ctx, ORF_SCOPE_MATERIAL, ground, "diffuse_color", half, 3));

        // MetallicMaterial::initPrograms (ORF/MetallicMaterial.cpp:108-118) and its variables
        int metal;
        orf_check(orf_material_create(ctx, "metallic_shader.cu", &metal));
        orf_check(orf_material_set_program(ctx, metal, ORF_RAY_RADIANCE, "metallic_shader.cu", "closest_hit"));
        orf_check(orf_material_set_program(ctx, metal, ORF_RAY_SHADOW, "metallic_shader.cu", "any_hit"));
        orf_check(orf_material_set_program(ctx, metal, ORF_RAY_DEPTH, "rough_depth_shader.cu", "closest_hit"));
        float rough[2] = {0.3f, 0.3f};
        orf_check(orf_variable_set_float(ctx, ORF_SCOPE_MATERIAL, metal, "roughness", rough, 2));
        uint32_t model = 1, dist = 1;
        orf_check(orf_variable_set_uint(ctx, ORF_SCOPE_MATERIAL, metal, "microfacet_model", &model, 1));
        orf_check(orf_variable_set_uint(ctx, ORF_SCOPE_MATERIAL, metal, "normal_distribution", &dist, 1));
        double ior[6] = {0.27, 2.78, 0.27, 2.78, 0.27, 2.78};  // MyComplex3
        orf_check(orf_variable_set_user_data(ctx, ORF_SCOPE_MATERIAL, metal, "ior", ior, sizeof(ior)));

        // RoughTranslucentMaterial::loadParameters (ORF/RoughTranslucentMaterial.cpp:360-364)
        int milk;
        orf_check(orf_material_create(ctx, "subsurface_scattering_shader.cu", &milk));
        const float sa[3] = {0.0011f, 0.0024f, 0.014f}, ss[3] = {2.55f, 3.21f, 3.77f}, g[3] = {0, 0, 0};
        orf_scattering_material_properties props = milk_properties(sa, ss, g, 100.0f, 1.3f);
        orf_check(orf_variable_set_user_data(ctx, ORF_SCOPE_MATERIAL, milk, "scattering_properties", &props,
                                             sizeof(props)));
        uint32_t dipole = 0;
        orf_check(orf_variable_set_uint(ctx, ORF_SCOPE_MATERIAL, milk, "dipole_model", &dipole, 1));
        uint32_t probes = 4096, per_hit = 4;
        orf_check(orf_variable_set_uint(ctx, ORF_SCOPE_CONTEXT, 0, "probe_count", &probes, 1));
        orf_check(orf_variable_set_uint(ctx, ORF_SCOPE_CONTEXT, 0, "probe_samples", &per_hit, 1));

        int emitter;
        orf_check(orf_material_create(ctx, "arealight_shader.cu", &emitter));

        // ORF/ObjLoader.cpp:198-410  geometry + geometry instance per mesh
        int gmesh, mmesh, smesh, lmesh;
        orf_check(orf_mesh_create(ctx, gv.data(), ngv, gt.data(), ngt, nullptr, 0, nullptr, ground, &gmesh));
        orf_check(orf_mesh_create(ctx, sv.data(), nsv, st.data(), nst, sn.data(), nsv, nullptr, metal, &mmesh));
        orf_check(orf_mesh_create(ctx, sv.data(), nsv, st.data(), nst, sn.data(), nsv, nullptr, milk, &smesh));
        orf_check(orf_mesh_create(ctx, lv.data(), 4, lt.data(), 2, nullptr, 0, nullptr, emitter, &lmesh));
        // ORF/OptixSceneLoader.cpp:497-498  gi["transform_matrix"] / ["inverse_transform_matrix"]
        const float to_left[16] = {1, 0, 0, -1.2f, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        const float from_left[16] = {1, 0, 0, 1.2f, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        orf_check(orf_variable_set_matrix4x4(ctx, ORF_SCOPE_INSTANCE, mmesh, "transform_matrix", 0, to_left));
        orf_check(orf_variable_set_matrix4x4(ctx, ORF_SCOPE_INSTANCE, mmesh, "inverse_transform_matrix", 0, from_left));
        const float to_right_colmajor[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 1.2f, 0, 0, 1};
        orf_check(orf_variable_set_matrix4x4(ctx, ORF_SCOPE_INSTANCE, smesh, "transform_matrix", 1, to_right_colmajor));
        // ORF/Light.cpp:242  gi["radiance"] of the light geometry
        float ten[3] = {10.0f, 10.0f, 10.0f};
        orf_check(orf_variable_set_float(ctx, ORF_SCOPE_INSTANCE, lmesh, "radiance", ten, 3));

        // ORF/OptixSceneLoader.cpp:297-347  triangle_light_buffer + light_buffer (RT_FORMAT_USER)
        std::vector<orf_triangle_light> tl(2);
        for (int t = 0; t < 2; ++t) {
            orf_triangle_light& x = tl[t];
            std::memset(&x, 0, sizeof(x));
            float* vv[3] = {x.v0, x.v1, x.v2};
            float* nn[3] = {x.n0, x.n1, x.n2};
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) {
                    vv[k][a] = lv[3 * lt[3 * t + k] + a];
                    nn[k][a] = a == 1 ? -1.0f : 0.0f;
                }
            x.has_normals = 1;
            for (int a = 0; a < 3; ++a) x.emission[a] = ten[a];
            // area = |(v1 - v2) x (v0 - v2)| / 2 (ORF/ObjLoader.cpp:532-612)
            float e1[3], e2[3];
            for (int a = 0; a < 3; ++a) {
                e1[a] = x.v1[a] - x.v2[a];
                e2[a] = x.v0[a] - x.v2[a];
            }
            const double cx = (double)e1[1] * e2[2] - (double)e1[2] * e2[1];
            const double cy = (double)e1[2] * e2[0] - (double)e1[0] * e2[2];
            const double cz = (double)e1[0] * e2[1] - (double)e1[1] * e2[0];
            x.area = (float)(0.5 * std::sqrt(cx * cx + cy * cy + cz * cz));
        }
        orf_check(orf_buffer_set(ctx, "triangle_light_buffer", tl.data(), sizeof(orf_triangle_light), tl.size()));
        orf_light_struct lights[2];
        std::memset(lights, 0, sizeof(lights));
        lights[0].light_type = 2;  // TRIANGLES_AREA_LIGHT {count, first, last}
        for (int a = 0; a < 3; ++a) lights[0].emitted_radiance[a] = ten[a];
        lights[0].light_parameters[0] = 2.0f;
        lights[0].light_parameters[1] = 0.0f;
        lights[0].light_parameters[2] = 1.0f;
        lights[1].light_type = 0;  // POINT_LIGHT
        const float prad[3] = {3.0f, 2.5f, 2.0f}, ppos[3] = {-2.0f, 2.5f, 1.5f};
        for (int a = 0; a < 3; ++a) {
            lights[1].emitted_radiance[a] = prad[a];
            lights[1].light_parameters[a] = ppos[a];
        }
        orf_check(orf_buffer_set(ctx, "light_buffer", lights, sizeof(orf_light_struct), 2));

        // ORF/Background.cpp:57-105 + ORF/HDRLoader.cpp:226-289  texture sampler "envmap"
        int tex;
        orf_check(orf_texture_create(ctx, env.data(), ew, eh, ORF_WRAP_REPEAT, ORF_FILTER_LINEAR, &tex));
        orf_check(orf_variable_set_texture(ctx, ORF_SCOPE_CONTEXT, 0, "envmap", tex));

        // a bad call raises, as optix::Exception did
        bool raised = false;
        try {
            float x = 1.0f;
            orf_check(orf_variable_set_float(ctx, ORF_SCOPE_MATERIAL, metal, "no_such_variable", &x, 1));
        } catch (const OrfException& e) {
            raised = e.code == ORF_ERROR_INVALID_VALUE && std::string(e.what()).find("no_such_variable") != std::string::npos;
        }
        if (!raised) throw std::runtime_error("an unknown variable was not refused");

        // ORF/ObjLoader.cpp:387-402 markDirty, ORF/OptixScene.cpp:139-159 launch
        orf_check(orf_accel_mark_dirty(ctx));
        int64_t stats[6];
        orf_check(orf_accel_build(ctx, stats));
        orf_check(orf_film_clear(ctx));
        orf_check(orf_launch(ctx, W, H, 0, (uint32_t)spp, nullptr));
        std::vector<double> film((size_t)W * H * 4);
        orf_check(orf_film_read_sums(ctx, film.data()));
        uint64_t cnt[4];
        orf_check(orf_counters(ctx, cnt));
        // ORF/OptixScene.cpp:68-75 destroyContext
        orf_check(orf_context_destroy(ctx));

        FILE* o = std::fopen(argv[2], "wb");
        if (!o || std::fwrite(film.data(), sizeof(double), film.size(), o) != film.size())
            throw std::runtime_error("cannot write film");
        std::fclose(o);
        std::printf("orf_caller: %dx%d spp %d, %lld triangles, %llu paths, %llu extension rays\n", W, H, spp,
                    (long long)stats[4], (unsigned long long)cnt[3], (unsigned long long)cnt[1]);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "orf_caller: %s\n", e.what());
        return 1;
    }
}
