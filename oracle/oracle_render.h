/* oracle_render.h -- CPU restatement of the per-pixel render path (TEST INFRASTRUCTURE).
 *
 * A scene is given as flattened world-space arrays; the oracle builds its own
 * acceleration structure (object-median BVH, independent of the product's
 * binned-SAH builder) and renders by following the reference's programs
 * recursively: CU/path_tracer.cu (ray generation), CU/diffuse_shader.cu,
 * CU/metallic_shader.cu, CU/arealight_shader.cu (closest hits),
 * CU/envmap_background.cu (miss), ORF/LightSampler.h, ORF/Envmap.h.
 */
#ifndef ORF_ORACLE_RENDER_H
#define ORF_ORACLE_RENDER_H
#include <stdint.h>

#include "orf_oracle.h"

enum {
    ORC_LAMBERT = 0,
    ORC_CONDUCTOR = 1,
    ORC_EMITTER = 2,
    ORC_DIELECTRIC = 3,
    ORC_INTERFACE = 4,
    ORC_TRANSLUCENT = 5,
    ORC_ROUGH_TRANSLUCENT = 6,
    ORC_TRANSPARENT = 7, /* transparent_shader.cu */
    ORC_NORMAL = 8,      /* normal_shader.cu */
    ORC_FLAT = 9,        /* flat_shader.cu */
    ORC_ROUGH_DIFFUSE = 10 /* rough_diffuse_shader.cu */
};
enum { ORC_POINT = 0, ORC_DIRECTIONAL = 1, ORC_TRIANGLES = 2, ORC_DISK = 3, ORC_SPHERE = 4, ORC_ENV = 5 };

typedef struct {
    int32_t kind;
    int32_t microfacet_model;
    int32_t normal_distribution;
    int32_t pad;
    float diffuse[3];
    float radiance[3];
    float absorption[3];
    float roughness[2];
    float ior;
    double ior_re[3], ior_im[3];
    /* translucent (ScatteringMaterial parameters, 1/mm, and scale) */
    float sss_absorption[3], sss_scattering[3], sss_meancosine[3];
    float sss_scale;
    int32_t dipole_model;
    int32_t pad2;
    /* flat material: bit 0 flat_shadow, bit 1 highlight; the highlight colour
     * is `radiance` */
    int32_t flat_flags;
    float flat_threshold;
} orc_material;

typedef struct {
    uint32_t type;
    float radiance[3];
    float params[3];  /* point: position; directional: direction; triangles: count, start, end;
                         disk, sphere: centre */
    float shape[3];   /* disk: radius, theta, phi (degrees); sphere: radius */
    int32_t material; /* disk, sphere: material of the emitter the rays see */
} orc_light;

typedef struct {
    float v[9];
    float n[9];
    uint32_t has_normals;
    float area;
} orc_trilight;

typedef struct {
    int64_t nverts, ntris, nnormals;
    const float* verts;
    const int32_t* idx;
    const float* normals;
    const int32_t* nidx;     /* NULL: geometric normals */
    const int32_t* tri_mat;
    int32_t nmats;
    const orc_material* mats;
    int32_t nlights;         /* the environment entry, if any, is appended by the oracle */
    const orc_light* lights;
    int32_t ntrilights;
    const orc_trilight* trilights;
    const float* env;        /* RGBA FP32, NULL = no environment */
    int32_t env_w, env_h, env_is;
    float eye[3], U[3], V[3], W[3];
    int32_t max_depth;
    float scene_epsilon;
} orc_scene_desc;

void* orc_scene_create(const orc_scene_desc* d);
void orc_scene_destroy(void* s);
/* Sum of samples frame0 .. frame0+spp-1 for pixels [x0,x1) x [y0,y1) of a
 * W x H film: out has 4 doubles per pixel (R, G, B sums of finite samples,
 * count); counters: [0] non-finite, [1] extension rays, [2] shadow rays, [3] paths. */
void orc_render(void* s, int W, int H, int x0, int y0, int x1, int y1, uint32_t frame0, uint32_t spp,
                double* out, uint64_t* counters, int threads);
/* Per-sample radiance of one frame for a pixel list: out has 3 doubles per pixel. */
void orc_render_samples(void* s, int W, int H, const int32_t* px, const int32_t* py, int n, uint32_t frame,
                        double* out, int threads);
/* The depth integrator's per-pixel (R, G, B, hit distance) of one frame. */
void orc_depth_samples(void* s, int W, int H, const int32_t* px, const int32_t* py, int n, uint32_t frame,
                       double* out);
/* The environment tables the oracle built (FP64 -> out arrays as double). */
void orc_env_tables(void* s, double* marg_pdf, double* marg_cdf, double* cond_pdf, double* cond_cdf);
void orc_env_sample_batch(void* s, uint32_t seed, int n, double* dir, double* pdf_sample, double* pdf_eval);
int orc_env_bsearch(const double* cdf, int n, double xi);
/* Closest hit of a batch of rays against the oracle BVH (t, prim) and by brute force. */
void orc_closest_hits(void* s, int n, const float* o, const float* d, float tmin, float* t, int32_t* prim,
                      int brute_force);
/* Translucent material (new estimator, see oracle_render.c): probes given by the
 * caller (4 floats per probe), nearest-probe search radius `cell`, gather samples
 * per entering hit; probe pass for sample index `frame` (outputs may be NULL);
 * probe values (L, w12 with T12 in w, and the entry normals sn, which may be NULL)
 * may also be loaded directly, 4 floats per probe each. */
void orc_probes_set(void* s, int n, const float* pos, const float* nrm, float cell, int samples, int material);
void orc_probe_pass(void* s, uint32_t frame, float* L, float* w12, float* sn, int threads);
void orc_probes_values(void* s, const float* L, const float* w12, const float* sn);
/* Derived coefficients of material m (64 floats, the product's field order). */
void orc_sss_coefficients(void* s, int m, float* out);
/* Standard / directional dipole of material m: standard at distances r[i];
 * directional at (xi, ni, w12, xo, no) per row of 15 floats.  out: 3 per entry. */
void orc_dipole_eval(void* s, int m, int n, const float* r, float* out);
void orc_dirpole_eval(void* s, int m, int n, const float* rows, float* out);
/* The directional dipole's infinite-medium term S_infinite alone: rows of
 * (r^2, x.w12, no.w12, x.no); out: 3 per row. */
void orc_dirpole_S_infinite_eval(void* s, int m, int n, const double* rows, double* out);
/* The rough translucent boundary's NEE reflection factor R(l.h) G1(i) G1(l) D_wi(h) /
 * (4 |i.n| |l.n|) for direction pairs about n = +z (3 floats each). */
void orc_rough_nee_reflect_batch(int n, const float* wi, const float* wl, float ax, float ay, int dist, float recip,
                                 double* out);
void orc_sss_gather_eval(void* s, int m, const float* xo, const float* no, int reps, uint32_t seed, double* out);

/* Differentiable form (new): textured albedo / roughness with frozen sampling,
 * and per-view evaluation of images and of the linear loss sum(dimg * mean L). */
void orc_diff_set(void* s, int tex_w, int tex_h, int alb_mat, int alb_uv, const double* alb_uvp, int rough_mat,
                  int rough_uv, const double* rough_uvp, const double* albedo_eval, const double* albedo_frozen,
                  const double* rough_eval, const double* rough_frozen);
/* Scattering coefficients of material m as parameters: shade with eval6, sample
 * (gather exit points, extinction roulette) with frozen6 (absorption RGB,
 * scattering RGB in 1/mm); m < 0 restores the materials' own. */
void orc_diff_set_sss(void* s, int m, const double* eval6, const double* frozen6);
/* The differentiable form's texel lookup for n points (xyz each): mode 0 planar
 * x-z over [uvp0, uvp2] x [uvp1, uvp3], mode 1 spherical about uvp[0..2]. */
void orc_diff_texel_batch(int n, int mode, const double* uvp, const double* pts, int tw, int th, int32_t* out);
double orc_diff_eval(void* s, int W, int H, int nviews, const float* cams, const uint32_t* frame0, uint32_t spp,
                     uint32_t shift, const double* dimg, double* images, int threads);

#endif
