// orf_sceneload.cpp -- JSON scene loader on the C-ABI (host).
//
// Restates OptixSceneLoader::readJSON (ORF/OptixSceneLoader.cpp:105-177) and
// the readJSON / loadFromJSON of the classes it dispatches to: PathTracer
// (ORF/Integrator.cpp:31-50), ConstantBackground / EnvMapBackground
// (ORF/Background.cpp:21-100, HDR through hdr_read), PinholeCamera
// (ORF/Camera.cpp:17-45, basis as calculateCameraVariables :83-103), OBJGeometry
// (ORF/Geometry.cpp:73-131; transform T * Rz * Ry * Rx * R(axis, angle) * S,
// :286-291), the material classes (mtl_name dispatch ORF/Geometry.cpp:186-216;
// parameter names of DiffuseMaterial, MetallicMaterial, LambertianInterface-,
// RoughTransparent-, Transparent-, Normal-, Flat-, RoughDiffuse-, Translucent-
// and RoughTranslucentMaterial::loadFromJSON)
// and the lights (ORF/Light.cpp: PointLight :22-40, DirectionalLight :75-93,
// TriangleAreaLight :149-290 with applyTransformationMatrix, DiskAreaLight
// :367-502, SphericalLight :507-607).  Every scene
// element becomes the same C-ABI calls a hand-written host would make.
//
// Differences: relative paths resolve against the JSON file's directory (the
// reference resolves them against its build directory); keys a file omits
// take this library's material defaults (the reference's JSON constructors
// leave them unset); AnisotropicMaterial ("anisotropic_structure_shader"), the
// one material class with no radiance program here, and integrators other than
// PathTracer and DepthTracer are refused with an error naming them.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "orf_assets.h"
#include "orf_common.h"
#include "orf_context.h"

namespace orf {
namespace {

struct Mat4 {
    double m[16];  // row-major
    static Mat4 identity()
    {
        Mat4 r;
        for (int i = 0; i < 16; ++i) r.m[i] = (i % 5 == 0) ? 1.0 : 0.0;
        return r;
    }
    Mat4 operator*(const Mat4& b) const
    {
        Mat4 r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double s = 0.0;
                for (int k = 0; k < 4; ++k) s += m[4 * i + k] * b.m[4 * k + j];
                r.m[4 * i + j] = s;
            }
        return r;
    }
};

Mat4 translate(const double t[3])
{
    Mat4 r = Mat4::identity();
    r.m[3] = t[0];
    r.m[7] = t[1];
    r.m[11] = t[2];
    return r;
}

Mat4 scale(const double s[3])
{
    Mat4 r = Mat4::identity();
    r.m[0] = s[0];
    r.m[5] = s[1];
    r.m[10] = s[2];
    return r;
}

// Rotation by `deg` degrees about the unit axis (right-handed, Rodrigues).
Mat4 rotate(double deg, double x, double y, double z)
{
    double len = std::sqrt(x * x + y * y + z * z);
    Mat4 r = Mat4::identity();
    if (len == 0.0) return r;
    x /= len;
    y /= len;
    z /= len;
    double a = deg * 3.14159265358979323846 / 180.0, c = std::cos(a), s = std::sin(a), t = 1.0 - c;
    r.m[0] = t * x * x + c;
    r.m[1] = t * x * y - s * z;
    r.m[2] = t * x * z + s * y;
    r.m[4] = t * x * y + s * z;
    r.m[5] = t * y * y + c;
    r.m[6] = t * y * z - s * x;
    r.m[8] = t * x * z - s * y;
    r.m[9] = t * y * z + s * x;
    r.m[10] = t * z * z + c;
    return r;
}

// 3x3 inverse transpose of the upper-left block, for normals.
void normal_matrix(const Mat4& M, double n[9])
{
    const double* a = M.m;
    double c00 = a[5] * a[10] - a[6] * a[9], c01 = a[6] * a[8] - a[4] * a[10], c02 = a[4] * a[9] - a[5] * a[8];
    double c10 = a[2] * a[9] - a[1] * a[10], c11 = a[0] * a[10] - a[2] * a[8], c12 = a[1] * a[8] - a[0] * a[9];
    double c20 = a[1] * a[6] - a[2] * a[5], c21 = a[2] * a[4] - a[0] * a[6], c22 = a[0] * a[5] - a[1] * a[4];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    ORF_REQUIRE(det != 0.0, "scene: singular transform");
    // inverse = cofactor^T / det; its transpose = cofactor / det
    double inv = 1.0 / det;
    double c[9] = {c00, c01, c02, c10, c11, c12, c20, c21, c22};
    for (int i = 0; i < 9; ++i) n[i] = c[i] * inv;
}

const JsonValue* params(const JsonValue& o, const char* key = "parameters")
{
    const JsonValue* p = o.get(key);
    return p && p->is_object() ? p : nullptr;
}

bool num(const JsonValue* p, const char* key, double& out)
{
    const JsonValue* v = p ? p->get(key) : nullptr;
    if (!v || !v->is_number()) return false;
    out = v->num;
    return true;
}

bool vec(const JsonValue* p, const char* key, double* out, size_t n)
{
    const JsonValue* v = p ? p->get(key) : nullptr;
    if (!v || !v->is_array()) return false;
    ORF_REQUIRE(v->arr.size() >= n, std::string("scene: '") + key + "' needs " + std::to_string(n) + " numbers");
    for (size_t i = 0; i < n; ++i) {
        ORF_REQUIRE(v->arr[i].is_number(), std::string("scene: '") + key + "' must hold numbers");
        out[i] = v->arr[i].num;
    }
    return true;
}

std::string str(const JsonValue& o, const char* key)
{
    const JsonValue* v = o.get(key);
    return v && v->is_string() ? v->str : std::string();
}

bool iequals(const std::string& a, const char* b)
{
    size_t n = std::strlen(b);
    if (a.size() != n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
    return true;
}

std::string resolve(const std::string& base, const std::string& p)
{
    if (p.empty() || p[0] == '/' || base.empty()) return p;
    return base + "/" + p;
}

void check(orf_status s)
{
    if (s != ORF_SUCCESS) throw std::runtime_error(orf_last_error());
}

// Object-to-world of an OBJ geometry or triangle light: position/scale,
// "rotation" [deg, ax, ay, az], x/y/z_rotation in degrees.
Mat4 object_transform(const JsonValue* p, const char* position_key)
{
    double t[3] = {0, 0, 0}, s[3] = {1, 1, 1}, r[4] = {0, 0, 1, 0}, a = 0;
    vec(p, position_key, t, 3);
    vec(p, "scale", s, 3);
    Mat4 R = Mat4::identity(), Rx = Mat4::identity(), Ry = Mat4::identity(), Rz = Mat4::identity();
    if (vec(p, "rotation", r, 4)) R = rotate(r[0], r[1], r[2], r[3]);
    if (num(p, "x_rotation", a)) Rx = rotate(a, 1, 0, 0);
    if (num(p, "y_rotation", a)) Ry = rotate(a, 0, 1, 0);
    if (num(p, "z_rotation", a)) Rz = rotate(a, 0, 0, 1);
    return translate(t) * (Rz * Ry * Rx * R) * scale(s);
}

void set_f(orf_context ctx, int scope, int h, const char* name, const double* v, int n)
{
    float f[4];
    for (int i = 0; i < n; ++i) f[i] = (float)v[i];
    check(orf_variable_set_float(ctx, scope, h, name, f, n));
}

void set_u(orf_context ctx, int scope, int h, const char* name, double v)
{
    uint32_t u = (uint32_t)v;
    check(orf_variable_set_uint(ctx, scope, h, name, &u, 1));
}

// A material from {"mtl_name": ..., "mtl_parameters": {...}}.
int load_material(orf_context ctx, const JsonValue& mj)
{
    const std::string name = str(mj, "mtl_name");
    const JsonValue* p = params(mj, "mtl_parameters");
    const char* program = nullptr;
    if (iequals(name, "diffuse_shader")) program = "diffuse_shader.cu";
    else if (iequals(name, "metallic_shader")) program = "metallic_shader.cu";
    else if (iequals(name, "lambertian_interface_shader")) program = "lambertian_interface_shader.cu";
    else if (iequals(name, "rough_transparent_shader")) program = "rough_transparent_shader.cu";
    else if (iequals(name, "translucent_shader")) program = "subsurface_scattering_shader.cu";
    else if (iequals(name, "rough_translucent_shader")) program = "rough_subsurface_scattering_shader.cu";
    else if (iequals(name, "transparent_shader")) program = "transparent_shader.cu";
    else if (iequals(name, "normal_shader")) program = "normal_shader.cu";
    else if (iequals(name, "flat_shader")) program = "flat_shader.cu";
    else if (iequals(name, "rough_diffuse_shader")) program = "rough_diffuse_shader.cu";
    else throw std::runtime_error("scene: material '" + name + "' has no radiance program here");
    int m = -1;
    check(orf_material_create(ctx, program, &m));
    const int S = ORF_SCOPE_MATERIAL;
    double v[3], x;
    if (vec(p, "diffuse_color", v, 3)) set_f(ctx, S, m, "diffuse_color", v, 3);
    if (iequals(name, "metallic_shader")) {
        double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
        bool hr = vec(p, "real_ior", re, 3), hi = vec(p, "imaginary_ior", im, 3);
        if (hr || hi) {
            double ior[6] = {re[0], im[0], re[1], im[1], re[2], im[2]};
            check(orf_variable_set_user_data(ctx, S, m, "ior", ior, sizeof(ior)));
        }
    } else if (num(p, "ior", x)) {
        set_f(ctx, S, m, "ior", &x, 1);
    }
    double r2[2];
    const JsonValue* rv = p ? p->get("roughness") : nullptr;
    if (rv && rv->is_array() && vec(p, "roughness", r2, rv->arr.size() >= 2 ? 2 : 1)) {
        if (rv->arr.size() < 2) r2[1] = r2[0];
        set_f(ctx, S, m, "roughness", r2, 2);
    } else if (num(p, "roughness", x)) {
        set_f(ctx, S, m, "roughness", &x, 1);
    }
    if (num(p, "microfacet_model", x)) set_u(ctx, S, m, "microfacet_model", x);
    if (num(p, "normal_distribution", x)) set_u(ctx, S, m, "normal_distribution", x);
    if (vec(p, "glass_absorption", v, 3)) set_f(ctx, S, m, "glass_absorption", v, 3);
    for (const char* k : {"absorption", "scattering", "meancosine"})
        if (vec(p, k, v, 3)) set_f(ctx, S, m, k, v, 3);
    if (num(p, "scale", x)) set_f(ctx, S, m, "scale", &x, 1);
    if (num(p, "dipole_model", x)) set_u(ctx, S, m, "dipole_model", x);
    // FlatMaterial::readJSON (ORF/FlatMaterial.cpp:205-245)
    if (vec(p, "highlight_color", v, 3)) set_f(ctx, S, m, "highlight_color", v, 3);
    if (num(p, "flat_shadow", x)) set_u(ctx, S, m, "flat_shadow", x);
    if (num(p, "highlight", x)) set_u(ctx, S, m, "highlight", x);
    if (num(p, "highlight_threshold", x)) set_f(ctx, S, m, "highlight_threshold", &x, 1);
    if (num(p, "shininess", x)) set_f(ctx, S, m, "shininess", &x, 1);
    return m;
}

// Transformed copy of an OBJ's triangles as triangle-light records (vertices by
// M, normals by its inverse transpose, as applyTransformationMatrix does).
void add_triangle_light(orf_context ctx, const ObjMesh& o, const Mat4& M, const double rad[3])
{
    double N[9];
    normal_matrix(M, N);
    const int64_t nt = (int64_t)(o.tris.size() / 3);
    std::vector<float> tv((size_t)nt * 9), tn;
    const bool has_n = !o.nidx.empty();
    if (has_n) tn.resize((size_t)nt * 9);
    for (int64_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) {
            const float* p = &o.verts[3 * (size_t)o.tris[3 * t + k]];
            for (int r = 0; r < 3; ++r)
                tv[9 * t + 3 * k + r] =
                    (float)(M.m[4 * r] * p[0] + M.m[4 * r + 1] * p[1] + M.m[4 * r + 2] * p[2] + M.m[4 * r + 3]);
            if (has_n) {
                const float* q = &o.normals[3 * (size_t)o.nidx[3 * t + k]];
                for (int r = 0; r < 3; ++r)
                    tn[9 * t + 3 * k + r] = (float)(N[3 * r] * q[0] + N[3 * r + 1] * q[1] + N[3 * r + 2] * q[2]);
            }
        }
    float fr[3] = {(float)rad[0], (float)rad[1], (float)rad[2]};
    check(orf_light_add_triangles(ctx, tv.data(), has_n ? tn.data() : nullptr, nt, fr));
}

int add_mesh(orf_context ctx, const ObjMesh& o, int material, const Mat4& M)
{
    int mesh = -1;
    const bool has_n = !o.nidx.empty();
    check(orf_mesh_create(ctx, o.verts.data(), (int64_t)(o.verts.size() / 3), o.tris.data(),
                          (int64_t)(o.tris.size() / 3), has_n ? o.normals.data() : nullptr,
                          has_n ? (int64_t)(o.normals.size() / 3) : 0, has_n ? o.nidx.data() : nullptr, material,
                          &mesh));
    float f[16];
    for (int i = 0; i < 16; ++i) f[i] = (float)M.m[i];
    check(orf_mesh_set_transform(ctx, mesh, f));
    return mesh;
}

}  // namespace

void scene_load_json(orf_context ctx, const std::string& path, uint32_t flags, orf_scene_info* info)
{
    std::string text;
    {
        FILE* f = std::fopen(path.c_str(), "rb");
        ORF_REQUIRE(f, "scene: cannot open '" + path + "'");
        char buf[65536];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
        std::fclose(f);
    }
    const JsonValue doc = json_parse(text);
    ctx->scene_doc = text;
    bool translucent_seen = false;
    ORF_REQUIRE(doc.is_object(), "scene: the document must be a JSON object");
    const size_t slash = path.find_last_of('/');
    const std::string base = slash == std::string::npos ? std::string() : path.substr(0, slash);
    orf_scene_info out;
    std::memset(&out, 0, sizeof(out));
    double x;
    if (num(&doc, "frame_count", x)) out.frame_count = (int32_t)x;
    const std::string bp = str(doc, "buffer_path");
    if (!bp.empty()) std::snprintf(out.buffer_path, sizeof(out.buffer_path), "%s", resolve(base, bp).c_str());

    if (const JsonValue* ij = doc.get("Integrator")) {
        const std::string type = str(*ij, "type");
        const bool depth = iequals(type, "DepthTracer");
        ORF_REQUIRE(iequals(type, "PathTracer") || depth, "scene: integrator '" + type + "' has no program here");
        check(orf_integrator_set(ctx, depth ? ORF_INTEGRATOR_DEPTH : ORF_INTEGRATOR_PATH_TRACER));
        const JsonValue* p = params(*ij);
        if (num(p, "max_depth", x)) {
            int32_t d = (int32_t)x;
            check(orf_variable_set_int(ctx, ORF_SCOPE_CONTEXT, 0, "max_depth", &d, 1));
        }
        if (num(p, "scene_epsilon", x)) set_f(ctx, ORF_SCOPE_CONTEXT, 0, "scene_epsilon", &x, 1);
    }

    if (const JsonValue* bj = doc.get("Background")) {
        const std::string type = str(*bj, "type");
        const JsonValue* p = params(*bj);
        double bg[3] = {0, 0, 0};
        vec(p, "background_color", bg, 3);
        const float solid[4] = {(float)bg[0], (float)bg[1], (float)bg[2], 1.0f};
        if (iequals(type, "EnvMapBackground")) {
            const std::string hp = str(*p, "path");
            HdrImage img;
            bool ok = !hp.empty();
            if (ok) {
                try {
                    img = hdr_read(resolve(base, hp));
                } catch (const Error&) {
                    ok = false;  // loadHDRTexture falls back to the background color (ORF/HDRLoader.cpp:243-259)
                } catch (const std::exception&) {
                    ok = false;
                }
            }
            if (ok)
                check(orf_envmap_set(ctx, img.rgba.data(), img.width, img.height,
                                     (flags & ORF_LOAD_ENV_IMPORTANCE) ? 1 : 0));
            else
                check(orf_envmap_set(ctx, solid, 1, 1, 0));
        } else if (iequals(type, "ConstantBackground")) {
            check(orf_envmap_set(ctx, solid, 1, 1, 0));
        } else {
            throw std::runtime_error("scene: background '" + type + "' has no program here");
        }
    }

    if (const JsonValue* cj = doc.get("Camera")) {
        const std::string type = str(*cj, "type");
        ORF_REQUIRE(iequals(type, "PinholeCamera"), "scene: camera '" + type + "' has no program here");
        const JsonValue* p = params(*cj);
        double eye[3] = {0, 0, 0}, at[3] = {0, 0, -1}, up[3] = {0, 1, 0}, vfov = 60, w = 512, h = 512;
        vec(p, "eye", eye, 3);
        vec(p, "lookat", at, 3);
        vec(p, "up", up, 3);
        num(p, "vfov", vfov);
        num(p, "width", w);
        num(p, "height", h);
        // calculateCameraVariables (ORF/Camera.cpp:83-103): W unnormalised
        double W[3] = {at[0] - eye[0], at[1] - eye[1], at[2] - eye[2]};
        double wlen = std::sqrt(W[0] * W[0] + W[1] * W[1] + W[2] * W[2]);
        double U[3] = {W[1] * up[2] - W[2] * up[1], W[2] * up[0] - W[0] * up[2], W[0] * up[1] - W[1] * up[0]};
        double ul = std::sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
        for (double& c : U) c /= ul;
        double V[3] = {U[1] * W[2] - U[2] * W[1], U[2] * W[0] - U[0] * W[2], U[0] * W[1] - U[1] * W[0]};
        double vl = std::sqrt(V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
        for (double& c : V) c /= vl;
        double vlen = wlen * std::tan(0.5 * vfov * 3.14159265358979323846 / 180.0);
        for (double& c : V) c *= vlen;
        for (double& c : U) c *= vlen * (w / h);
        set_f(ctx, ORF_SCOPE_CONTEXT, 0, "eye", eye, 3);
        set_f(ctx, ORF_SCOPE_CONTEXT, 0, "U", U, 3);
        set_f(ctx, ORF_SCOPE_CONTEXT, 0, "V", V, 3);
        set_f(ctx, ORF_SCOPE_CONTEXT, 0, "W", W, 3);
        out.width = (int32_t)w;
        out.height = (int32_t)h;
    }

    if (const JsonValue* ga = doc.get("Geometries")) {
        ORF_REQUIRE(ga->is_array(), "scene: 'Geometries' must be an array");
        for (const JsonValue& gj : ga->arr) {
            const std::string type = str(gj, "type");
            ORF_REQUIRE(iequals(type, "obj"), "scene: geometry '" + type + "' has no program here");
            const JsonValue* p = params(gj);
            ORF_REQUIRE(p && !str(*p, "path").empty(), "scene: an obj geometry needs a path");
            ObjMesh o = obj_read(resolve(base, str(*p, "path")));
            const JsonValue* mj = p->get("material");
            int m;
            if (mj && mj->is_object()) {
                m = load_material(ctx, *mj);
                const std::string mn = str(*mj, "mtl_name");
                translucent_seen = translucent_seen || iequals(mn, "translucent_shader") ||
                                   iequals(mn, "rough_translucent_shader");
            } else {  // OBJGeometry's default DiffuseMaterial
                check(orf_material_create(ctx, "diffuse_shader.cu", &m));
            }
            add_mesh(ctx, o, m, object_transform(p, "position"));
            out.meshes++;
            out.materials++;
        }
    }

    if (const JsonValue* la = doc.get("Lights")) {
        ORF_REQUIRE(la->is_array(), "scene: 'Lights' must be an array");
        for (const JsonValue& lj : la->arr) {
            const std::string type = str(lj, "type");
            const JsonValue* p = params(lj);
            double rad[3] = {1, 1, 1}, v[3] = {0, 0, 0};
            vec(p, "radiance", rad, 3);
            float fr[3] = {(float)rad[0], (float)rad[1], (float)rad[2]};
            if (iequals(type, "PointLight")) {
                vec(p, "position", v, 3);
                float fp[3] = {(float)v[0], (float)v[1], (float)v[2]};
                check(orf_light_add_point(ctx, fp, fr));
            } else if (iequals(type, "DirectionalLight")) {
                vec(p, "direction", v, 3);
                float fd[3] = {(float)v[0], (float)v[1], (float)v[2]};
                check(orf_light_add_directional(ctx, fd, fr));
            } else if (iequals(type, "DiskAreaLight") || iequals(type, "SphericalLight")) {
                // DiskAreaLight / SphericalLight::readJSON (ORF/Light.cpp:389-423, 525-548):
                // defaults radius 1 (disk) / 0.05 (sphere), theta = phi = 0
                const bool disk = iequals(type, "DiskAreaLight");
                double radius = disk ? 1.0 : 0.05, theta = 0.0, phi = 0.0;
                vec(p, "position", v, 3);
                num(p, "radius", radius);
                num(p, "theta", theta);
                num(p, "phi", phi);
                float fp[3] = {(float)v[0], (float)v[1], (float)v[2]};
                int m = -1;
                if (disk)
                    check(orf_light_add_disk(ctx, fp, (float)radius, (float)theta, (float)phi, fr, &m));
                else
                    check(orf_light_add_sphere(ctx, fp, (float)radius, fr, &m));
                out.materials++;
            } else if (iequals(type, "TrianglesAreaLight")) {
                ORF_REQUIRE(p && !str(*p, "path").empty(), "scene: a triangle light needs a path");
                ObjMesh o = obj_read(resolve(base, str(*p, "path")));
                const Mat4 M = object_transform(p, "position");
                int m = -1;
                check(orf_material_create(ctx, "arealight_shader.cu", &m));
                check(orf_variable_set_float(ctx, ORF_SCOPE_MATERIAL, m, "radiance", fr, 3));
                add_mesh(ctx, o, m, M);
                add_triangle_light(ctx, o, M, rad);
                out.meshes++;
                out.materials++;
            } else {
                throw std::runtime_error("scene: light '" + type + "' has no sampler here");
            }
            out.lights++;
        }
    }
    // probe set of the first translucent object (the reference's per-frame probe
    // buffer, ORF/OptixSceneLoader.cpp:11, 482-522): 2^16 Poisson-disk probes
    // unless the caller sets probe_count itself
    if (translucent_seen) {
        uint32_t probes = 1u << 16;
        check(orf_variable_set_uint(ctx, ORF_SCOPE_CONTEXT, 0, "probe_count", &probes, 1));
    }
    if (info) *info = out;
}

}  // namespace orf

extern "C" orf_status orf_scene_load_json(orf_context ctx, const char* path, uint32_t flags, orf_scene_info* info)
{
    ORF_API_BEGIN
    ORF_REQUIRE(ctx && path, "orf_scene_load_json: bad arguments");
    orf::scene_load_json(ctx, path, flags, info);
    ORF_API_END
}

extern "C" orf_status orf_hdr_read(const char* path, int32_t* width, int32_t* height, float* rgba)
{
    ORF_API_BEGIN
    ORF_REQUIRE(path && width && height, "orf_hdr_read: bad arguments");
    orf::HdrImage img = orf::hdr_read(path);
    *width = img.width;
    *height = img.height;
    if (rgba) std::memcpy(rgba, img.rgba.data(), img.rgba.size() * sizeof(float));
    ORF_API_END
}

extern "C" orf_status orf_obj_read(const char* path, int64_t* nverts, int64_t* ntris, int64_t* nnormals,
                                   int32_t* has_nindex, float* verts, int32_t* vindex, float* normals,
                                   int32_t* nindex)
{
    ORF_API_BEGIN
    ORF_REQUIRE(path && nverts && ntris && nnormals && has_nindex, "orf_obj_read: bad arguments");
    orf::ObjMesh o = orf::obj_read(path);
    *nverts = (int64_t)(o.verts.size() / 3);
    *ntris = (int64_t)(o.tris.size() / 3);
    *nnormals = (int64_t)(o.normals.size() / 3);
    *has_nindex = o.nidx.empty() ? 0 : 1;
    if (verts) std::memcpy(verts, o.verts.data(), o.verts.size() * sizeof(float));
    if (vindex) std::memcpy(vindex, o.tris.data(), o.tris.size() * sizeof(int32_t));
    if (normals) std::memcpy(normals, o.normals.data(), o.normals.size() * sizeof(float));
    if (nindex && !o.nidx.empty()) std::memcpy(nindex, o.nidx.data(), o.nidx.size() * sizeof(int32_t));
    ORF_API_END
}
