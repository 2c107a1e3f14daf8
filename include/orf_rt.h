/*
 * orf_rt.h -- C-ABI of the MI355X-native render runtime (liborf_rt.so).
 *
 * This is the drop-in boundary that replaces the OptiX 5 host runtime the
 * reference's scene classes call (optixu/optixpp_namespace.h as used from
 * ORF/OptixScene.cpp, ORF/OptixSceneLoader.cpp, ORF/...Material.cpp,
 * ORF/Light.cpp, ORF/Camera.cpp, ORF/Integrator.cpp, ORF/Background.cpp,
 * ORF/ObjLoader.cpp).  Plain C: opaque handles, plain pointers and sizes, one
 * status code per call and orf_last_error() for the message (the role of
 * RTresult + optix::Exception).  Device pointers are HIP device addresses;
 * `stream` is a hipStream_t (NULL = the default stream).
 *
 * Threading: one host thread owns a context (as one Qt thread owned the
 * optix::Context); calls on different contexts may run concurrently.
 */
#ifndef ORF_RT_H
#define ORF_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum orf_status {
    ORF_SUCCESS = 0,
    ORF_ERROR_INVALID_VALUE = 1,  /* bad argument / handle / variable name  */
    ORF_ERROR_HIP = 2,            /* HIP runtime failure                     */
    ORF_ERROR_OUT_OF_MEMORY = 3,  /* host or device allocation failed        */
    ORF_ERROR_NOT_FOUND = 4,      /* file / program / variable not found     */
    ORF_ERROR_FILE_FORMAT = 5,    /* malformed OBJ / HDR / JSON / raw film   */
    ORF_ERROR_UNKNOWN = 99
} orf_status;

/* Text of the last error raised on the calling thread ("" if none). */
const char* orf_last_error(void);
/* Library version string, e.g. "orf_rt 0.1 gfx950". */
const char* orf_version(void);

/* ------------------------------------------------------------------------
 * Rough-conductor BSDF operator (batch of independent queries).
 *
 * Replaces the visible-normal sampling step of the Metallic material's
 * closest-hit program, CU/metallic_shader.cu:84-158 (microfacet normal from
 * ORF/Microfacet.h:72-83 -> ORF/MicrofacetGGX.h:190-207, masking
 * ORF/MicrofacetGGX.h:63-74, Fresnel ORF/Fresnel.h:78-91), with the material
 * parameters MetallicMaterial sets (ior, roughness: ORF/MetallicMaterial.cpp).
 * There, one query ran per radiance hit inside OptiX traversal; here it is a
 * flat SoA batch in local shading space (macro-normal +z).
 *
 * Inputs (device, FP32, n each): wo (towards the viewer, unit, z > 0),
 * isotropic GGX roughness alpha, uniform numbers u1 u2.  ior_re/ior_im
 * (host, 3 each): complex IOR per RGB channel, relative to air (n1 = 1).
 * Outputs (device, FP32, n each): wi = sampled reflection direction,
 * pdf = D_wo(m) / (4 |wi.m|) in solid angle, weight = F(wo.m) G1(wi)
 * (= f cos / pdf).  weight is 0 where the reference stops the path.
 * ------------------------------------------------------------------------ */
orf_status orf_bsdf_ggx_conductor_batch(int64_t n,
                                        const float* wo_x, const float* wo_y, const float* wo_z,
                                        const float* alpha, const float* u1, const float* u2,
                                        const double ior_re[3], const double ior_im[3],
                                        float* wi_x, float* wi_y, float* wi_z, float* pdf,
                                        float* w_r, float* w_g, float* w_b,
                                        void* stream);

/* Host-side generator of the synthetic operator workload: one PCG32 stream
 * (O'Neill 2014, initstate seed, stream seq), five draws per query:
 * cos(theta)=u0, phi=2 pi u1, alpha=0.05+0.95 u2, u1=u3, u2=u4.  Host arrays. */
orf_status orf_gen_bsdf_queries(int64_t n, uint64_t seed, uint64_t seq,
                                float* wo_x, float* wo_y, float* wo_z, float* alpha,
                                float* u1, float* u2);

/* Inspection of the translucent programs' 64-bit random stream, evaluated on
 * the host by the same code the kernels run: n draws of rnd_accurate
 * (ORF/random.h:134-142, MD5 of ORF/md5.h) from the state (x, y), with the
 * state after each draw (2 words per draw); and the value step alone,
 * (float)((double)l / 2^63) for l < 2^63.  No counterpart in the reference. */
orf_status orf_rng_accurate_stream(uint32_t x, uint32_t y, int n, uint32_t* states, float* values);
orf_status orf_rng_unit63(const uint64_t* l, int n, float* out);

/* ------------------------------------------------------------------------
 * Render context: scene objects, named variables, launches, film.
 *
 * Mirrors the subset of the optix:: host API the reference's scene classes
 * use.  A context owns one device, one scene and one FP64 film.
 * ------------------------------------------------------------------------ */
typedef struct orf_context_t* orf_context;

/* optix::Context::create + setRayTypeCount/setEntryPointCount/setStackSize
 * (ORF/OptixScene.cpp:96-117).  device = HIP ordinal. */
orf_status orf_context_create(int device, orf_context* out);
/* OptixScene::destroyContext (ORF/OptixScene.cpp:68-75). */
orf_status orf_context_destroy(orf_context ctx);

/* Variable scopes of context["name"] / material["name"] / geometry["name"] /
 * geometry_instance["name"].  A mesh is one Geometry inside one
 * GeometryInstance here, so ORF_SCOPE_GEOMETRY and ORF_SCOPE_INSTANCE both
 * take a mesh handle and address that instance: a material variable set there
 * overrides the material's value for that mesh only (OptiX looks instance
 * variables up before material ones), as the reference sets gi["radiance"] on
 * light geometry (ORF/Light.cpp:242, 419-422) and gi["scattering_properties"]
 * on translucent instances (ORF/OptixSceneLoader.cpp:497-501). */
enum { ORF_SCOPE_CONTEXT = 0, ORF_SCOPE_MATERIAL = 1, ORF_SCOPE_GEOMETRY = 2, ORF_SCOPE_INSTANCE = 3 };

/* ScatteringMaterialProperties (ORF/structs.h:116-144) byte for byte: 65 floats,
 * 260 bytes.  Set as user data "scattering_properties" on a translucent
 * material (RoughTranslucentMaterial::loadParameters, ORF/RoughTranslucentMaterial.cpp:
 * 360-364) or instance; the shading then reads these fields as given, derived
 * ones included, as the reference's programs read the struct.  ior_real and
 * ior_imag are carried and unused, as in the reference's programs. */
typedef struct orf_scattering_material_properties {
    float relative_ior;
    float ior_real[3], ior_imag[3];
    float absorption[3], scattering[3], meancosine[3];
    float extinction[3], reducedExtinction[3], deltaEddExtinction[3], D[3], transport[3];
    float C_phi[3], C_phi_inv[3], C_E[3], A[3], albedo[3], reducedAlbedo[3], de[3], three_D[3], two_A_de[3];
    float global_coeff[3], one_over_three_ext[3];
    float mean_transport;
} orf_scattering_material_properties;

/* LightStruct (ORF/structs.h:42-48, 40 bytes) and TriangleLight (:95-102, 92
 * bytes), the element types of the context buffers "light_buffer" and
 * "triangle_light_buffer" (ORF/OptixSceneLoader.cpp:297-347). */
typedef struct orf_light_struct {
    uint32_t light_type;  /* 0 point, 1 directional, 2 triangles, 3 disk, 4 sphere (ORF/structs.h:146-153) */
    float emitted_radiance[3];
    float light_parameters[3];  /* point: position; directional: direction; triangles: count, first, last;
                                   disk, sphere: position */
    float padding[3];           /* disk: radius, phi, theta (degrees); sphere: radius */
} orf_light_struct;
typedef struct orf_triangle_light {
    float v0[3], v1[3], v2[3];
    float n0[3], n1[3], n2[3];
    uint32_t has_normals;
    float emission[3];
    float area;
} orf_triangle_light;

/* Variable setters: context["name"]->setFloat/setInt/setUint(...) and
 * material["name"]->... with 1-4 components; setUserData(size, ptr) for
 * structured values (e.g. the conductor "ior" as MyComplex3, 6 doubles:
 * ORF/MetallicMaterial.cpp, ORF/MyComplex.h:7-16).  Context variables:
 * max_depth, scene_epsilon, eye, U, V, W, frame, patch_origin, patch_dims,
 * exception_color (ORF/Integrator.cpp:6-24, ORF/Camera.cpp:83-103).
 * Material variables: diffuse_color, radiance, roughness, ior,
 * microfacet_model, normal_distribution, glass_absorption; translucent:
 * absorption, scattering, meancosine, scale, dipole_model; flat: flat_shadow,
 * highlight, highlight_color, highlight_threshold, shininess. */
orf_status orf_variable_set_float(orf_context ctx, int scope, int handle, const char* name,
                                  const float* v, int n);
orf_status orf_variable_set_int(orf_context ctx, int scope, int handle, const char* name,
                                const int32_t* v, int n);
orf_status orf_variable_set_uint(orf_context ctx, int scope, int handle, const char* name,
                                 const uint32_t* v, int n);
orf_status orf_variable_set_user_data(orf_context ctx, int scope, int handle, const char* name,
                                      const void* data, size_t size);
/* setMatrix4x4fv(transpose, m) (ORF/OptixSceneLoader.cpp:497-498, 557-558):
 * row-major 4x4 unless transpose != 0.  Instance / geometry scope:
 * "transform_matrix" is the mesh's object-to-world transform (as
 * orf_mesh_set_transform), "inverse_transform_matrix" must be its inverse
 * (checked to 1e-4, else ORF_ERROR_INVALID_VALUE).  Context scope:
 * "transform_matrix" and "normal_matrix" of the probe program's current object
 * are accepted and have no effect (the probe pass takes every translucent
 * mesh's own transform). */
orf_status orf_variable_set_matrix4x4(orf_context ctx, int scope, int handle, const char* name, int transpose,
                                      const float m[16]);
/* createBuffer(RT_BUFFER_INPUT, RT_FORMAT_USER) + setElementSize + map/copy/unmap
 * + context[name]->set(buffer) for the light buffers
 * (ORF/OptixSceneLoader.cpp:297-347): "light_buffer" (element_size 40,
 * orf_light_struct) replaces the light list, its disk and sphere entries
 * bringing their emitters as orf_light_add_disk / _sphere do;
 * "triangle_light_buffer" (element_size 92, orf_triangle_light) replaces the
 * triangles the triangle entries index.  Either may be set first; a triangle
 * entry indexing past the triangle buffer is refused at the next launch. */
orf_status orf_buffer_set(orf_context ctx, const char* name, const void* data, size_t element_size, size_t count);

/* createMaterial + setClosestHitProgram(radiance, createProgramFromPTXFile(file,
 * "closest_hit")) (e.g. ORF/MetallicMaterial.cpp:108-118).  `program` is the
 * reference program file: "diffuse_shader.cu", "metallic_shader.cu",
 * "arealight_shader.cu", "rough_transparent_shader.cu",
 * "lambertian_interface_shader.cu", "subsurface_scattering_shader.cu"
 * (also "translucent_shader"), "rough_subsurface_scattering_shader.cu",
 * "transparent_shader.cu", "normal_shader.cu", "flat_shader.cu",
 * "rough_diffuse_shader.cu".  Returns the material handle. */
orf_status orf_material_create(orf_context ctx, const char* program, int* material);
/* setClosestHitProgram / setAnyHitProgram(ray_type, createProgramFromPTXFile(file,
 * entry)) (ORF/MetallicMaterial.cpp:108-118).  ray_type 0 radiance: entry
 * "closest_hit" of any program orf_material_create accepts selects the
 * material's shading (its variables are kept); ray_type 1 shadow: entry
 * "any_hit" -- every reference any-hit is opaque except normal_shader.cu's,
 * which lets shadow rays through; ray_type 2 depth: "closest_hit" of
 * "depth_shader.cu" or "rough_depth_shader.cu" (the depth integrator answers
 * every material with the hit distance).  Other combinations are refused with
 * ORF_ERROR_NOT_FOUND. */
enum { ORF_RAY_RADIANCE = 0, ORF_RAY_SHADOW = 1, ORF_RAY_DEPTH = 2 };
orf_status orf_material_set_program(orf_context ctx, int material, int ray_type, const char* program_file,
                                    const char* entry);

/* createGeometry + vertex/index/normal buffers + createGeometryInstance
 * (ORF/ObjLoader.cpp:198-410).  verts: 3 floats per vertex; vindex: 3 ints per
 * triangle; normals/nindex optional (NULL: geometric normals; nindex NULL with
 * normals given: nindex = vindex).  One material per mesh. */
orf_status orf_mesh_create(orf_context ctx, const float* verts, int64_t nverts, const int32_t* vindex,
                           int64_t ntris, const float* normals, int64_t nnormals, const int32_t* nindex,
                           int material, int* mesh);
/* Transform node: setMatrix(transpose=false, m) with a row-major object-to-world
 * 4x4 (ORF/Geometry.cpp:286-291); normals use its inverse transpose. */
orf_status orf_mesh_set_transform(orf_context ctx, int mesh, const float m[16]);

/* LightStruct entries of light_buffer (ORF/OptixSceneLoader.cpp:349-410). */
orf_status orf_light_add_point(orf_context ctx, const float position[3], const float radiance[3]);
orf_status orf_light_add_directional(orf_context ctx, const float direction[3], const float radiance[3]);
/* TrianglesAreaLight: appends one TriangleLight per triangle to
 * triangle_light_buffer (area = |(v1-v2) x (v0-v2)| / 2, ORF/ObjLoader.cpp:532-612)
 * and one LightStruct {count, start, end}.  tri_verts: 9 floats per triangle in
 * world space; tri_normals: 9 floats per triangle or NULL. */
orf_status orf_light_add_triangles(orf_context ctx, const float* tri_verts, const float* tri_normals,
                                   int64_t ntris, const float radiance[3]);
/* DiskAreaLight (ORF/Light.cpp:367-505): a one-sided disk emitter of the given
 * radius whose normal is (sin theta sin phi, cos theta, sin theta cos phi),
 * theta and phi in degrees; visible to rays as the disk intersection of
 * CU/disk_mesh.cu:41-63 with an arealight material, sampled for NEE by
 * evaluate_disk_area_light (ORF/LightSampler.h:73-101).  The emitter's
 * material handle is returned in *emitter_material (may be NULL); batch
 * remaps must cover it like any mesh material. */
orf_status orf_light_add_disk(orf_context ctx, const float position[3], float radius, float theta, float phi,
                              const float radiance[3], int* emitter_material);
/* SphericalLight (ORF/Light.cpp:507-607): a sphere emitter (CU/sphere_mesh.cu:42-91)
 * sampled by evaluate_spherical_area_light (ORF/LightSampler.h:103-141). */
orf_status orf_light_add_sphere(orf_context ctx, const float position[3], float radius, const float radiance[3],
                                int* emitter_material);
/* Removes every light, the disk and sphere emitters included. */
orf_status orf_lights_clear(orf_context ctx);

/* EnvMapBackground (ORF/Background.cpp:57-105): FP32 RGBA lat-long texture,
 * row 0 at the nadir.  importance_sample != 0 builds the FP64 2D CDF on the GPU
 * (spec: CU/env_cameras.cu) and adds the environment to the light list. */
orf_status orf_envmap_set(orf_context ctx, const float* rgba, int width, int height, int importance_sample);
/* createTextureSampler + setWrapMode/setFilteringModes + setBuffer(FLOAT4)
 * (ORF/HDRLoader.cpp:226-289, ORF/Texture.h:18-84): an FP32 RGBA texture
 * (host array, row 0 first).  Returns the texture handle. */
enum { ORF_WRAP_REPEAT = 0, ORF_WRAP_CLAMP_TO_EDGE = 1 };
enum { ORF_FILTER_LINEAR = 0, ORF_FILTER_NEAREST = 1 };
orf_status orf_texture_create(orf_context ctx, const float* rgba, int width, int height, int wrap_mode,
                              int filter_mode, int* texture);
/* context[name]->setTextureSampler(texture): "envmap" (ORF/Background.cpp:57-105)
 * makes the texture the environment (REPEAT wrap and LINEAR filtering, as
 * loadHDRTexture creates it; others are refused), importance sampled unless
 * the context variable "envmap_importance" is 0. */
orf_status orf_variable_set_texture(orf_context ctx, int scope, int handle, const char* name, int texture);
/* Read back the FP32 pdf/cdf tables (host arrays: h, h, h*w, h*w). */
orf_status orf_envmap_tables(orf_context ctx, float* marg_pdf, float* marg_cdf, float* cond_pdf,
                             float* cond_cdf);

/* Geometry changed: BVH rebuilt at the next launch (Acceleration::markDirty). */
orf_status orf_accel_mark_dirty(orf_context ctx);
/* Build/upload now (optional; launch does it when dirty).  stats: nodes, leaves,
 * max_depth, max_leaf, tris, degenerate (6 int64, may be NULL). */
orf_status orf_accel_build(orf_context ctx, int64_t* stats);

/* Integrator pass over a width x height film (ORF/OptixScene.cpp:139-159 with
 * path_tracer, CU/path_tracer.cu:32-74): samples frame0 .. frame0+spp-1 of every
 * pixel in one launch, sample s seeded tea<16>(width*y + x, s).  The film keeps
 * FP64 sums of finite samples and their count; it is cleared when the size
 * changes or by orf_film_clear. */
orf_status orf_launch(orf_context ctx, int width, int height, uint32_t frame0, uint32_t spp, void* stream);
/* setRayGenerationProgram of the Integrator classes (ORF/Integrator.cpp:26-50
 * PathTracer, :100-118 DepthTracer): which program orf_launch runs.  The depth
 * tracer writes (1 + n) / 2 of the face-forwarded shading normal as RGB and
 * the hit distance as alpha (CU/depth_tracer.cu:33-58, CU/depth_shader.cu:13-19);
 * orf_film_read returns that alpha.  Batch and differentiable renders always
 * path trace. */
enum { ORF_INTEGRATOR_PATH_TRACER = 0, ORF_INTEGRATOR_DEPTH = 1 };
orf_status orf_integrator_set(orf_context ctx, int integrator);
orf_status orf_film_clear(orf_context ctx);
/* Film mean as RGBA (A = 1) in FP64 (rgba64) and/or FP32 (rgba32); host arrays
 * of 4*width*height, row 0 = bottom row (launch index y = 0). */
orf_status orf_film_read(orf_context ctx, double* rgba64, float* rgba32);
/* Display buffer: the film mean as orf_film_read gives it (RGBA, FP32, same
 * layout) written into a DEVICE array of 4*width*height floats on `stream`,
 * without synchronising.  Replaces the reference's output buffer created on a GL
 * pixel-buffer object (createBufferFromGLBO, ORF/OptixScene.cpp:252-257) that
 * OptixWindow::paintBuffer uploads to an RGBA32F texture
 * (ORF/OptixWindow.cpp:146-196): a GL host registers its PBO with HIP-GL
 * interop, maps it and passes the mapped pointer here once per frame. */
orf_status orf_film_display(orf_context ctx, float* rgba32_device, void* stream);
/* Raw film (sum R, G, B, count) as FP64 host array of 4*width*height. */
orf_status orf_film_read_sums(orf_context ctx, double* sums);
/* Counters since the last clear: [0] non-finite samples dropped,
 * [1] extension (closest-hit) rays, [2] shadow rays, [3] paths. */
orf_status orf_counters(orf_context ctx, uint64_t out[4]);
/* No counterpart in the reference: orf_launch for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same kernels, seeds and film arithmetic as orf_launch
 * (path tracer integrator only), but the film is the caller's: `film` is a
 * device array of 4*width*height doubles (sum R, G, B, count per pixel, row 0 =
 * bottom row) that the launch adds to.  `counters` (device, 4 uint64, may be
 * NULL) receives the orf_counters statistics the same way.  The scene must
 * already be on the device (orf_accel_build, or a previous launch, after the
 * last scene change), else ORF_ERROR_INVALID_VALUE.  The launch's own working
 * memory is a fixed 128-byte per-device static array (8 work-queue counters and
 * 4 fallback statistics counters), cleared by a kernel at the start of each
 * call: the call allocates nothing and needs no size limit beyond orf_launch's,
 * but calls that overlap on one device race on that array. */
orf_status orf_launch_noalloc(orf_context ctx, int width, int height, uint32_t frame0, uint32_t spp, double* film,
                              uint64_t* counters, void* stream);

/* ------------------------------------------------------------------------
 * Batched renders: many uneven renders of the context's scene in one call.
 *
 * No counterpart in the reference, which launches one integrator pass per
 * frame (ORF/OptixScene.cpp:139-159); each item here is that pass for one
 * camera/film, sample s of item i seeded tea<16>(width*y + x, frame0 + s).
 * remap (n x remap_stride ints, NULL = identity) gives each item its own
 * material table: geometry material id m shades as material remap[i][m].
 * ------------------------------------------------------------------------ */
typedef struct orf_render_item {
    int32_t width, height;
    uint32_t spp, frame0;
    float eye[3], U[3], V[3], W[3];
} orf_render_item;

/* Wavefront parameters (the autotuner's search space). */
typedef struct orf_wavefront_config {
    int32_t block_size;     /* traversal workgroup size: 64, 128 or 256 */
    int32_t pool_size;      /* path slots in flight (multiple of 256) */
    int32_t sort_policy;    /* 1: per-material shade queues, 0: one shade kernel */
    float regen_threshold;  /* refill free slots when continuing paths < threshold * pool */
    int32_t use_graph;      /* 1: bounce loop replayed as a hipGraph */
    int32_t graph_pairs;    /* bounce-iteration pairs per graph replay */
} orf_wavefront_config;

enum { ORF_BATCH_WAVEFRONT = 0, ORF_BATCH_MEGAKERNEL = 1, ORF_BATCH_PERSISTENT = 2 };

orf_status orf_wavefront_default_config(orf_wavefront_config* cfg);
/* mode ORF_BATCH_WAVEFRONT: all items in one wavefront pass; ORF_BATCH_MEGAKERNEL:
 * one megakernel launch per item; ORF_BATCH_PERSISTENT: one persistent megakernel
 * launch over every item (work items of 8x8 pixels x up to 8 samples, per-XCD
 * queues with stealing).  cfg (wavefront only) NULL = defaults.  *seconds (may be
 * NULL) receives the wall time from the first launch to completion. */
orf_status orf_render_batch(orf_context ctx, const orf_render_item* items, int32_t n, const int32_t* remap,
                            int32_t remap_stride, int32_t mode, const orf_wavefront_config* cfg, void* stream,
                            double* seconds);
/* Mean RGBA (A = 1) of item `item` of the last batch, FP64, 4*width*height, row 0 = bottom. */
orf_status orf_batch_film_read(orf_context ctx, int32_t item, double* rgba);
/* Raw FP64 sums (R, G, B, count) of item `item`. */
orf_status orf_batch_film_sums(orf_context ctx, int32_t item, double* sums);
/* [0] non-finite samples, [1] extension rays, [2] shadow rays, [3] paths of the last batch. */
orf_status orf_batch_counters(orf_context ctx, uint64_t out[4]);
/* Coordinate search over block size, pool size, sort policy and refill threshold,
 * timing the wavefront on the batch with every spp divided by spp_divisor.
 * Writes the fastest configuration and a JSON report of all trials. */
orf_status orf_wavefront_autotune(orf_context ctx, const orf_render_item* items, int32_t n, const int32_t* remap,
                                  int32_t remap_stride, uint32_t spp_divisor, orf_wavefront_config* best,
                                  char* report, size_t report_size);

/* ------------------------------------------------------------------------
 * Differentiable renders (no counterpart in the reference).
 *
 * Scenes with disk or sphere lights are refused.
 * A Lambert material's albedo and a GGX conductor's roughness can come from
 * textures (nearest texel of a planar x-z or a spherical mapping of the hit
 * point).  orf_diff_render writes per-view mean images; orf_diff_adjoint runs
 * the path-replay adjoint: given dLoss/dImage it adds dLoss/dTexel into the
 * gradient buffers (device, caller-zeroed) using samples decorrelated from the
 * primal's (frame index + frame_shift).  Sampling and Russian roulette are
 * detached (the BSDF weight is differentiated at the sampled direction).
 * Textures, images and gradients are device FP32 arrays:
 * albedo tex_h*tex_w*3, roughness tex_h*tex_w, images nviews*h*w*3.
 * ------------------------------------------------------------------------ */
typedef struct orf_diff_desc {
    int32_t albedo_material;      /* material handle (diffuse_shader) or -1 */
    int32_t albedo_uv;            /* 0: planar x-z, 1: spherical */
    float albedo_uv_params[4];    /* planar: x0, z0, x1, z1; spherical: centre x, y, z, unused */
    int32_t roughness_material;   /* material handle (metallic_shader, GGX) or -1 */
    int32_t roughness_uv;
    float roughness_uv_params[4];
    int32_t tex_w, tex_h;
    int32_t scattering_material;  /* translucent material handle whose coefficients are parameters, or -1 */
} orf_diff_desc;

orf_status orf_diff_render(orf_context ctx, const orf_diff_desc* desc, const orf_render_item* views, int32_t nviews,
                           const float* albedo, const float* roughness, float* images, void* stream);
orf_status orf_diff_adjoint(orf_context ctx, const orf_diff_desc* desc, const orf_render_item* views,
                            int32_t nviews, const float* albedo, const float* roughness, const float* d_images,
                            float* g_albedo, float* g_roughness, uint32_t frame_shift, void* stream);
/* The same with the scattering coefficients of desc->scattering_material as a
 * third parameter (a translucent material, subsurface_scattering_shader or
 * rough_subsurface_scattering_shader, whose coefficients come from its
 * variables, not from user data): `scattering` is a device array of 6 floats,
 * absorption RGB then scattering RGB in 1/mm (the material variables; its
 * scale, meancosine, ior and dipole model stay the material's).  The render
 * shades the material with them (the probe pass included); the adjoint adds
 * dLoss/d(absorption RGB, scattering RGB) into g_scattering (device, 6 floats,
 * caller-zeroed): the extinction weights inside the medium are differentiated
 * with their roulette probability held fixed, and the diffusion gather with its
 * exit points, their pdf and the probes held fixed.  orf_diff_render and
 * orf_diff_adjoint are these with scattering = NULL and refuse a desc whose
 * scattering_material is set. */
orf_status orf_diff_render_ex(orf_context ctx, const orf_diff_desc* desc, const orf_render_item* views, int32_t nviews,
                              const float* albedo, const float* roughness, const float* scattering, float* images,
                              void* stream);
orf_status orf_diff_adjoint_ex(orf_context ctx, const orf_diff_desc* desc, const orf_render_item* views,
                               int32_t nviews, const float* albedo, const float* roughness, const float* scattering,
                               const float* d_images, float* g_albedo, float* g_roughness, float* g_scattering,
                               uint32_t frame_shift, void* stream);
/* Statistics of the last differentiable launch (render or adjoint), after a
 * device synchronisation: out[0] = samples dropped as non-finite, out[1] =
 * adjoint samples whose replay ended on a radiance other than their first
 * pass's.  The replay re-runs the first pass's path on the same random numbers,
 * so out[1] is zero unless the two passes took different paths -- the gradients
 * of such a launch are not to be trusted. */
orf_status orf_diff_counters(orf_context ctx, uint64_t out[2]);

/* ------------------------------------------------------------------------
 * Translucent materials ("subsurface_scattering_shader.cu", smooth boundary;
 * "rough_subsurface_scattering_shader.cu", microfacet boundary with the
 * material variables roughness, microfacet_model and normal_distribution as
 * RoughTranslucentMaterial sets them, ORF/RoughTranslucentMaterial.cpp:14-28;
 * defaults Beckmann, visible normals, roughness 0.1).
 *
 * Material variables absorption, scattering, meancosine (3 floats, 1/mm), scale,
 * ior, dipole_model (0 standard, 1 directional) as ScatteringMaterial /
 * RoughTranslucentMaterial take them (ORF/RoughTranslucentMaterial.cpp:158-215);
 * defaults are the Wholemilk preset with scale 100.  Context variables
 * probe_count (0: no diffusion term), probe_samples (exit-point samples per
 * entering hit, default 4) and probe_seed select the Poisson-disk surface probes
 * that replace the reference's per-frame probe buffer (CU/sample_camera.cu);
 * every launch first runs the probe pass with sample index frame0.
 * ------------------------------------------------------------------------ */
/* Derived coefficients (ScatteringMaterialProperties, ORF/structs.h:116-144, as
 * ORF/RoughTranslucentMaterial.cpp:366-395 computes them): n >= 64 floats in the
 * order relative_ior, dipole_model, then the 3-vectors absorption, scattering,
 * meancosine, extinction, reducedExtinction, deltaEddExtinction, D, transport,
 * C_phi, C_phi_inv, C_E, A, albedo, reducedAlbedo, de, three_D, two_A_de,
 * global_coeff, one_over_three_ext, then mean_transport. */
orf_status orf_scattering_properties(orf_context ctx, int material, float* out, int n);
/* The probes (no reference counterpart).  Arrays of 4 floats per probe (xyz, w);
 * any pointer may be NULL.  L: incident radiance estimate of the last probe pass;
 * w12: refracted incident direction (xyz) and T12 (w; times the microfacet
 * boundary's transmission weight for "rough_subsurface_scattering_shader.cu");
 * sn: the normal the light entered through (the probe normal, or the sampled
 * microfacet normal of a rough boundary; PositionSample::normal after
 * CU/sample_camera.cu:130-151). */
orf_status orf_probes_read(orf_context ctx, int64_t* count, double* area, float* cell, float* rmin, float* pos,
                           float* nrm, float* L, float* w12, float* sn);
/* Run the probe pass alone for sample index `frame`: replaces the reference's
 * per-frame launch of the sample_camera ray-generation program
 * (ORF/OptixSceneLoader.cpp:152-153, CU/sample_camera.cu:48-268). */
orf_status orf_probe_pass(orf_context ctx, uint32_t frame, void* stream);

/* Procedural scene pieces (host).  Icosphere: call with verts = tris = normals =
 * NULL to get the sizes.  Env: sun_dir in env coordinates, sigma in degrees. */
orf_status orf_gen_icosphere(int level, uint32_t noise_seed, float amplitude, float frequency,
                             float* verts, float* normals, int64_t* nverts, int32_t* tris, int64_t* ntris);
orf_status orf_gen_sky_envmap(int width, int height, const float sun_dir[3], float sun_sigma_deg,
                              float sun_peak, float* rgba);

/* ------------------------------------------------------------------------
 * Scene files.
 *
 * orf_scene_load_json replaces OptixSceneLoader::readJSON
 * (ORF/OptixSceneLoader.cpp:105-177): Integrator (PathTracer, DepthTracer), Background
 * (EnvMapBackground with a Radiance .hdr, ConstantBackground), Camera
 * (PinholeCamera), Geometries (obj with a material {mtl_name, mtl_parameters})
 * and Lights (PointLight, DirectionalLight, TrianglesAreaLight, DiskAreaLight,
 * SphericalLight) become the calls above.  Relative paths resolve against the JSON file's directory.  The
 * reference does not importance-sample its environment; flag
 * ORF_LOAD_ENV_IMPORTANCE turns that on.  Classes without a program here are
 * refused with ORF_ERROR_INVALID_VALUE naming them.
 * ------------------------------------------------------------------------ */
enum { ORF_LOAD_ENV_IMPORTANCE = 1 };

typedef struct {
    int32_t width, height;   /* PinholeCamera width / height (0: no camera) */
    int32_t frame_count;     /* "frame_count" of the file (0 if absent) */
    int32_t meshes, materials, lights;
    char buffer_path[512];   /* "buffer_path", resolved; "" if absent */
} orf_scene_info;

orf_status orf_scene_load_json(orf_context ctx, const char* path, uint32_t flags, orf_scene_info* info);

/* ------------------------------------------------------------------------
 * Progressive frames and film files.
 *
 * The film accumulates FP64 sums and counts; its mean equals the reference's
 * FP32 running-mean output_buffer (CU/path_tracer.cu:64-68).  A .raw film is
 * width x height FP32 RGBA, row 0 = launch row 0 (the bottom of the image),
 * alpha 1, no header (OptixScene::saveScene, ORF/OptixScene.cpp:202-236).
 * ------------------------------------------------------------------------ */
/* OptixScene::renderScene for frames [frame_begin, frame_end)
 * (ORF/OptixScene.cpp:139-168): per frame, the translucent probe pass and one
 * sample per pixel with that frame's seeds (= orf_launch(frame, 1) per frame). */
orf_status orf_render_frames(orf_context ctx, int width, int height, uint32_t frame_begin, uint32_t frame_end,
                             void* stream);
/* The film's mean as a .raw file. */
orf_status orf_film_save_raw(orf_context ctx, const char* path);
/* OptixScene::loadBuffer (ORF/OptixScene.cpp:251-299): a width x height .raw
 * becomes the film after `frames` frames (sums = mean * frames, counts =
 * frames), so frames rendered next continue its running mean.  A file of
 * another size is refused with ORF_ERROR_FILE_FORMAT. */
orf_status orf_film_load_raw(orf_context ctx, const char* path, int width, int height, uint32_t frames);
/* bufferToImage + QImage::save (ORF/ImageLoader.cpp:110-133): 8-bit RGBA PNG of
 * the film's mean, top row first, each channel (int)(v * 255) clamped. */
orf_status orf_film_save_png(orf_context ctx, const char* path);
/* OptixScene::saveScene + OptixSceneLoader::saveJSONScene
 * (ORF/OptixSceneLoader.cpp:48-70): for path "<base>.<ext>" writes <base>.json
 * (the document last loaded by orf_scene_load_json with frame_count and
 * buffer_path = "<base name>.raw", relative to the document), <base>.png and
 * <base>.raw. */
orf_status orf_scene_save(orf_context ctx, const char* path, uint32_t frame_count);

/* HDRLoader + loadHDRTexture (ORF/HDRLoader.cpp:63-181, 226-289): Radiance RGBE,
 * -Y +X only; RGBA FP32 with row 0 = the file's bottom row.  rgba may be NULL
 * (size query). */
orf_status orf_hdr_read(const char* path, int32_t* width, int32_t* height, float* rgba);

/* glmReadOBJ + ObjLoader::loadVertexData / createGeometryInstances
 * (ORF/glm.cpp, ORF/ObjLoader.cpp:198-330): v, vn and f (polygons as fans,
 * negative indices relative); 0-based indices.  has_nindex = 0 when some face
 * lacks normals.  Array pointers may be NULL (size query). */
orf_status orf_obj_read(const char* path, int64_t* nverts, int64_t* ntris, int64_t* nnormals, int32_t* has_nindex,
                        float* verts, int32_t* vindex, float* normals, int32_t* nindex);

#ifdef __cplusplus
}
#endif

#endif /* ORF_RT_H */
