/*
 * rtg.h - C ABI of the MI355X-native MIS path tracer (librtg.so).
 *
 * The drop-in boundary replaces the reference renderer's per-job CPU loop
 *   Scene::compute -> JobManager::addJobs -> SampleJob::run -> PathTracingMIS::li
 *   (RaytracerNew2/src/core/Scene.cpp:83-101, core/JobManager.cpp:103-232,
 *    core/SampleJob.cpp:28-72, integrator/PathTracingMIS.cpp:27-263)
 * at image/job granularity: the host describes a scene once (meshes, spheres,
 * materials, lights), then asks for whole renders or submits render requests.
 *
 * Conventions: plain C types only; every call returns RTG_OK (0) or a negative
 * error code and never throws; rtg_last_error() gives the message of the last
 * failing call made by the calling thread. Host pointers unless a name says _device.
 * Primitive ids: triangles of mesh 0, mesh 1, ... in index order, then spheres.
 */
#ifndef RTG_H
#define RTG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RTG_OK = 0,
  RTG_ERR_INVALID = -1,    /* bad argument or unsupported feature */
  RTG_ERR_HIP = -2,        /* HIP runtime failure */
  RTG_ERR_OOM = -3,        /* device memory exhausted / over budget */
  RTG_ERR_PARSE = -4,      /* scene file could not be parsed */
  RTG_ERR_NOT_READY = -5,  /* job still running (rtg_poll) */
  RTG_ERR_CANCELLED = -6   /* job was cancelled */
};

/* Material types: bsdf/Diffuse.cpp, bsdf/RoughConductor.cpp, bsdf/RoughDielectric.cpp */
/* Diffuse (bsdf/Diffuse.cpp), RoughConductor (bsdf/RoughConductor.cpp), RoughDielectric
 * (bsdf/RoughDielectric.cpp), and the smooth (delta) Conductor (bsdf/Conductor.cpp:33-77) and
 * Dielectric (bsdf/Dielectric.cpp:28-110). */
enum {
  RTG_MAT_DIFFUSE = 0,
  RTG_MAT_ROUGH_CONDUCTOR = 1,
  RTG_MAT_ROUGH_DIELECTRIC = 2,
  RTG_MAT_CONDUCTOR = 3,
  RTG_MAT_DIELECTRIC = 4
};
/* Microfacet distributions: bsdf/MicrofacetDistribution.h:8-12 */
enum { RTG_DIST_BECKMANN = 0, RTG_DIST_GGX = 2 };
/* Light selection strategies: integrator/Integrator.h:18-22. RTG_LIGHTS_DEFAULT (0, what a
 * zero-initialised description holds) is the reference's default, one_light_uniform
 * (Integrator.cpp:14). */
enum { RTG_LIGHTS_DEFAULT = 0, RTG_LIGHTS_UNIFORM = 1, RTG_LIGHTS_WEIGHTED = 2, RTG_LIGHTS_ALL = 3 };
/* Samplers: sampler/RandomSampler.cpp, sampler/StratifiedSampler.cpp */
enum { RTG_FILTER_BOX = 0, RTG_FILTER_GAUSSIAN = 1, RTG_FILTER_MITCHELL = 2 };
enum { RTG_SAMPLER_RANDOM = 0, RTG_SAMPLER_STRATIFIED = 1 };
/* core/main.cpp:185-290 tone mappers */
enum { RTG_TONEMAP_CLAMP = 0, RTG_TONEMAP_GAMMA = 1, RTG_TONEMAP_REINHARD = 2 };
/* integrator/PathTracingMIS.cpp, integrator/VolPathTracing.cpp */
enum { RTG_INTEGRATOR_PATH_MIS = 0, RTG_INTEGRATOR_VOLPATH = 1 };
/* rtg_render_desc.guiding (no reference counterpart): RTG_GUIDING_OFF is the reference's estimator;
 * RTG_GUIDING_LEARNED adds the online-trained vMF-mixture guide at diffuse vertices of the first
 * three path depths and the rough-dielectric exit proposals; RTG_GUIDING_EXIT the exit proposals alone. Both guided modes
 * keep the plain estimator's expectation. */
enum { RTG_GUIDING_OFF = 0, RTG_GUIDING_LEARNED = 1, RTG_GUIDING_EXIT = 2 };

/* Constant-texture material parameters (texture/ConstantTexture.h:6-37).
 * albedo: Diffuse albedo, or reflectanceTexture of the rough BSDFs.
 * alpha: roughness as written in the scene (the BSDF uses Color(alpha).average()).
 * eta/k: conductor IOR and absorption (RoughConductor.cpp:24).
 * eta_int/eta_ext: dielectric IORs (RoughDielectric.cpp:24-25).
 * transmitted: RoughDielectric transmittedTexture. */
typedef struct rtg_material {
  int32_t type;
  int32_t distribution;
  float albedo[3];
  float alpha;
  float eta[3];
  float k[3];
  float eta_int;
  float eta_ext;
  float transmitted[3];
  /* Image textures (texture/ImageTexture.cpp, ScaleTexture.cpp) in place of the constant values
   * above: 0 = the constant, k >= 1 = rtg_scene_desc.textures[k - 1], evaluated at the hit's uv.
   * tex_albedo replaces albedo (Diffuse "albedo"; "reflectanceTexture" of the others),
   * tex_alpha the roughness (the texture's channel average, as the reference evaluates it),
   * tex_transmitted the dielectrics' "transmittedTexture". */
  int32_t tex_albedo;
  int32_t tex_alpha;
  int32_t tex_transmitted;
} rtg_material;

/* An ImageTexture (texture/ImageTexture.cpp:7-64): the decoded linear RGB texel array that the
 * reference's ImageLoader produces, looked up at uv as ImageTexture::eval does (x = u (width - 1),
 * y = v height, times uv_scale; optional inversion; wrapped; bilinear, core/BasicMath.h:74-90),
 * times value_scale (a ScaleTexture around it, texture/ScaleTexture.cpp:19-22; 1 without). */
typedef struct rtg_texture_desc {
  const float* rgb;            /* height x width x 3, row 0 first */
  int32_t width;
  int32_t height;
  float uv_scale;              /* ImageTexture "scale" (its absolute value is used) */
  int32_t invert_x;            /* ImageTexture "invertX" */
  int32_t invert_y;            /* ImageTexture "invertY" */
  float value_scale;           /* ScaleTexture "scale" */
} rtg_texture_desc;

/* A triangle mesh (geometricShape/Mesh.h:19-58). emission != NULL makes the
 * mesh a MeshLight with that radiance (light/MeshLight.cpp:32-64). */
typedef struct rtg_mesh_desc {
  const float* vertices;       /* num_vertices x 3 */
  const float* normals;        /* num_vertices x 3 or NULL */
  const float* uvs;            /* num_vertices x 2 or NULL */
  int32_t num_vertices;
  const int32_t* indices;      /* num_triangles x 3 */
  int32_t num_triangles;
  const int32_t* tri_material; /* per-triangle material ids or NULL */
  int32_t material;            /* material when tri_material is NULL */
  const float* emission;       /* 3 floats or NULL */
} rtg_mesh_desc;

/* geometricShape/Sphere.cpp:10-17: radius and toWorld (row-major 4x4). */
typedef struct rtg_sphere_desc {
  float to_world[16];
  float radius;
  int32_t material;
} rtg_sphere_desc;

/* Whole scene. Light order (for the selection CDF): emissive meshes in mesh
 * order, then the environment light if any (light/EnvironmentLight.cpp). */
typedef struct rtg_scene_desc {
  const rtg_material* materials;
  int32_t num_materials;
  const rtg_mesh_desc* meshes;
  int32_t num_meshes;
  const rtg_sphere_desc* spheres;
  int32_t num_spheres;
  const float* env_rgb;        /* env_height x env_width x 3 (lat-long) or NULL */
  int32_t env_width;
  int32_t env_height;
  float env_scale;
  int32_t light_strategy;      /* RTG_LIGHTS_DEFAULT (= UNIFORM), RTG_LIGHTS_UNIFORM, RTG_LIGHTS_WEIGHTED or
                                  RTG_LIGHTS_ALL (at most 8 lights; not served) */
  /* Camera medium (the scene file's top-level <medium>, core/XMLLoader.cpp:88-92,159-172): a
   * Heterogeneous medium (medium/Heterogeneous.cpp) over a GridVolume density grid
   * (volume/GridVolume.cpp) with a Henyey-Greenstein phase function (medium/HGPhaseFunction.cpp).
   * As a camera medium it is never placed in an owner box, so the grid spans the world box
   * [0,1]^3. Only the RTG_INTEGRATOR_VOLPATH integrator reads it. NULL density = no medium. */
  const float* medium_density; /* res_x * res_y * res_z values, x fastest then y then z */
  int32_t medium_res[3];       /* each >= 2 */
  float medium_albedo[3];      /* single-scattering albedo ("albedo") */
  float medium_g;              /* HGPhaseFunction "g" (< 1e-4: isotropic sampling) */
  float medium_scale;          /* density multiplier ("scale"; <= 0 takes the default 1) */
  const rtg_texture_desc* textures;  /* referenced by rtg_material.tex_* (1-based) */
  int32_t num_textures;
  /* 1: the camera medium is Homogeneous (medium/Homogeneous.cpp: "sigmaA", "sigmaS", "scale" as
   * medium_scale; unbounded; medium_density must be NULL; medium_g as above) */
  int32_t medium_homogeneous;
  float medium_sigma_a[3];
  float medium_sigma_s[3];
} rtg_scene_desc;

/* Camera (camera/NoriPerspective.cpp:10-37 + core/Transform.cpp:199-227 lookAt),
 * sampler and integrator parameters (integrator/PathTracingMIS.cpp:12-20). */
typedef struct rtg_render_desc {
  float origin[3];
  float target[3];
  float up[3];
  float fov;
  float near_clip;
  float far_clip;
  int32_t width;
  int32_t height;
  int32_t spp;
  int32_t sampler;             /* RTG_SAMPLER_* */
  int32_t dimension;           /* stratified dimensions (StratifiedSampler.cpp:24) */
  int32_t min_depth;
  int32_t max_depth;
  uint64_t seed;
  int32_t guiding;             /* RTG_GUIDING_* (no reference counterpart). OFF: plain MIS path
                                  tracing. LEARNED: a vMF-mixture network trained online on this
                                  render's own paths guides diffuse vertices at depths 0-2
                                  (one-sample MIS with BSDF sampling; rtg_render /
                                  rtg_render_device only), plus EXIT's proposals. It trains on at
                                  most 2^21 path slots (128 steps or 1/16 of the samples); on a
                                  context with a larger pool the rest of the render then runs on
                                  that pool with the trained network. EXIT: at
                                  rough-dielectric vertices on the denser side, one-sample MIS of
                                  the sampler with two proposals for the normals near the critical
                                  angle and near grazing, where the sampler's weights are unbounded
                                  (also served). PathTracingMIS only. */
  int32_t use_camera_matrix;   /* 1: camera_to_world (row-major, the XML camera's toWorld, e.g. a
                                  lookAt with a mirroring scale) replaces origin/target/up */
  float camera_to_world[16];
  /* Reconstruction filter (the camera's <filter>; filter/BoxFilter.h, GaussianFilter.cpp,
   * MitchellNetravaliFilter.cpp): RTG_FILTER_BOX (weight 1), RTG_FILTER_GAUSSIAN (param = alpha,
   * omega; a value <= 0 takes the reference's default, 2 and radius x) or RTG_FILTER_MITCHELL
   * (param = B, C as given; the reference's defaults are 1/3, 1/3), radius x/y (<= 0: 0.5; at
   * most 8), weights looked up in the reference's 32x32 table (ReconstructionFilter.h:53-62). A sample is splatted into the pixels of its
   * 32x32 block's sub-screen (block plus the filter's overlap border, Screen.cpp:44-111) at their
   * true positions; contributions outside the image are dropped. (The reference's final image
   * reads its bordered array from index 0, which shifts filtered images by the overlap; the film
   * returned here is not shifted.) All zero = the box filter of radius 0.5. */
  int32_t filter;
  float filter_radius[2];
  float filter_param[2];
  /* RTG_INTEGRATOR_PATH_MIS (0, integrator/PathTracingMIS.cpp) or RTG_INTEGRATOR_VOLPATH
   * (integrator/VolPathTracing.cpp:25-426: delta tracking in the scene's camera medium, ratio
   * tracking for transmittance; rtg_render / rtg_render_device only, no guiding) */
  int32_t integrator;
} rtg_render_desc;

typedef struct rtg_stats {
  uint64_t samples;            /* camera paths finished */
  uint64_t extension_rays;
  uint64_t shadow_rays;
  uint64_t nan_samples;        /* dropped as in core/Screen.cpp:46-50 */
  uint64_t iterations;         /* wavefront iterations */
  double milliseconds;         /* device time of the render */
  /* traversal work, filled when RTG_PROFILE_COUNT is set (rtg_ctx_set_profiling) */
  uint64_t nodes_visited;      /* interior nodes (64-byte child pairs) fetched by extension rays */
  uint64_t prims_tested;       /* primitives intersected by extension rays */
  uint64_t shadow_nodes_visited;
  uint64_t shadow_prims_tested;
  /* device time per stage and launches, filled when RTG_PROFILE_EVENTS is set:
   * 0 regen, 1 extend (+ classify), 2 guide inference, 3 shade diffuse, 4 shade conductor,
   * 5 shade dielectric, 6 miss, 7 shadow (+ compaction) */
  double stage_ms[8];
  uint64_t stage_launches[8];
  /* guided renders */
  uint64_t guide_train_steps;  /* training steps run during the render */
  uint64_t guide_samples;      /* training samples the render produced */
  double guide_loss;           /* loss of the last training step */
  double guide_train_ms;       /* device time of the training steps (RTG_PROFILE_EVENTS) */
  uint64_t guide_queries;      /* network inferences (guided vertices) */
  uint64_t stack_spills;       /* traversal-stack pushes past the 12 on-chip (LDS) entries into the
                                  global spill area (RTG_PROFILE_COUNT) */
} rtg_stats;

enum { RTG_PROFILE_EVENTS = 1, RTG_PROFILE_COUNT = 2 };

typedef struct rtg_ctx rtg_ctx;
typedef struct rtg_scene rtg_scene;

/* Context: one per GPU. Replaces JobManager's thread pool (core/JobManager.cpp:140-232). */
int rtg_ctx_create(int device, rtg_ctx** out);
int rtg_ctx_destroy(rtg_ctx* ctx);
const char* rtg_last_error(rtg_ctx* ctx);
/* Wavefront pool size (path slots); 0 restores the default of 2^21 slots (rtg_autotune picks one
 * per scene). */
int rtg_ctx_set_pool(rtg_ctx* ctx, int32_t slots);
/* Launch configuration of the wavefront renderer (no reference counterpart: the reference's
 * schedule is its JobManager thread count, core/JobManager.cpp:11). */
typedef struct rtg_tuning {
  int32_t pool_slots;    /* path slots of the wavefront pool (rounded up to a multiple of 256) */
  int32_t trace_waves;   /* waves per SIMD of the persistent traversal launches (1..8) */
  int32_t trace_refill;  /* idle lanes of a traversal wave before it takes new rays (1..64) */
  int32_t graphs;        /* 1: steady-state iteration batches replay captured hipGraphs */
} rtg_tuning;
int rtg_ctx_get_tuning(rtg_ctx* ctx, rtg_tuning* out);
/* Fails with RTG_ERR_INVALID while a server exists on the context. */
int rtg_ctx_set_tuning(rtg_ctx* ctx, const rtg_tuning* tuning);
/* Autotuner: times renders of `desc` on `scene` (its size, about trial_samples samples each,
 * 0 = 2^24; no guiding) under a coordinate search over the pool size (2^20 to 2^26 slots, at most
 * twice the trial's samples), then traversal waves and refill threshold, starting from the context's current
 * configuration; applies the fastest to the context and returns it (and its samples/s). Trials as
 * large as the real render let it find the pool that render wants. Fails with RTG_ERR_INVALID while
 * a server exists. */
int rtg_autotune(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc* desc, uint64_t trial_samples,
                 rtg_tuning* best, double* best_samples_per_s);
/* Profiling flags (RTG_PROFILE_*): per-stage HIP events and traversal work counters. */
int rtg_ctx_set_profiling(rtg_ctx* ctx, int32_t flags);

/* Scene upload: builds the BVH on the host and makes the scene resident.
 * Replaces Scene::initialize (core/Scene.cpp:115-138) + Bvh::build (accelStruct/Bvh.cpp:51-79). */
int rtg_scene_upload(rtg_ctx* ctx, const rtg_scene_desc* desc, rtg_scene** out);
/* Scene file in the reference's XML schema (core/XMLLoader.cpp:33-904): the
 * hot-path types only. Fills *render from <camera>, <sampler>, <integrator>. */
int rtg_scene_load_xml(rtg_ctx* ctx, const char* path, rtg_scene** out, rtg_render_desc* render);
/* A sampled spectrum (n >= 2 (lambda nm, value) pairs) -> linear sRGB, as the scene loader turns a
 * named metal's IOR spectra (data/IOR/<metal>.eta.spd, .k.spd) into the conductor's eta and k.
 * Replaces Color3::setFromSampledSpectrum (core/Color.cpp:423-443) as called from
 * IORHelper::extractAndConvertSpectrumFromFile (tools/Common.cpp:83-110): CIE 1931 2-degree
 * integration at 1 nm over 360-830 nm, negative channels clamped to zero. Host only. */
int rtg_spectrum_to_rgb(const float* lambda, const float* values, int32_t n, float* out_rgb);
int rtg_scene_release(rtg_ctx* ctx, rtg_scene* scene);
/* Device bytes held by a scene. */
int rtg_scene_bytes(rtg_scene* scene, uint64_t* bytes);
/* What a resident scene holds (for loaders and tests). */
typedef struct rtg_scene_info {
  int32_t triangles, spheres, materials, lights, env_width, env_height;
  float bbox_min[3], bbox_max[3];
  /* camera medium (all 0 without one): grid resolution, albedo, phase g, density scale and the
   * tracking majorant (maximum density x scale, Heterogeneous.cpp:40) */
  int32_t medium_res[3];
  float medium_albedo[3];
  float medium_g, medium_scale, medium_sigma_max;
  int32_t textures;            /* image textures */
  int32_t medium_kind;         /* 0 none, 1 heterogeneous grid, 2 homogeneous */
  float medium_sigma_a[3], medium_sigma_s[3];  /* homogeneous, after the scale */
  int32_t traversal_stack;     /* deepest traversal stack a ray can need (the BVH builder's bound) */
} rtg_scene_info;
int rtg_scene_query(rtg_scene* scene, rtg_scene_info* info);
/* Material `index` of a resident scene as uploaded (for a scene file: as the loader built it, named
 * metals converted to eta/k). No reference counterpart (introspection for loaders and tests). */
int rtg_scene_material(rtg_scene* scene, int32_t index, rtg_material* out);

/* Blocking render. out_rgbw (width*height*4 floats, host) receives per-pixel
 * sums of radiance and the filter weight (box filter r=0.5); divide rgb by w
 * to get the image (core/Screen.cpp:113-122). */
int rtg_render(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc* desc, float* out_rgbw, rtg_stats* stats);
/* Same with a caller-owned device buffer (width*height*4 floats), not cleared. */
int rtg_render_device(rtg_ctx* ctx, rtg_scene* scene, const rtg_render_desc* desc, float* out_rgbw_device,
                      rtg_stats* stats);

/* The image the reference displays and saves (core/main.cpp:185-290, after Pixel::applyColor,
 * core/Screen.cpp:8-15): rgbw (width*height*4 floats, host, as rtg_render returns it) is resolved
 * (rgb / w where w != 0), then RTG_TONEMAP_CLAMP (x 255), RTG_TONEMAP_GAMMA (times
 * 2^((exposure - 0.5) * 20), then the sRGB curve; the reference's default exposure is 0.5) or
 * RTG_TONEMAP_REINHARD (global: L' = L (1 + L / Lmax^2) / (1 + L) on the luminance, then sRGB),
 * clamped to [0, 255] and truncated to 8 bits -> out_rgb (width*height*3 bytes, host). */
int rtg_tonemap(rtg_ctx* ctx, const float* rgbw, int32_t width, int32_t height, int32_t mode, float exposure,
                uint8_t* out_rgb);

/* Parity tap for Scene::computeIntersection (core/Scene.cpp:106-113, accelStruct/Bvh.cpp:392-481).
 * rays: n x 8 floats (origin, direction, minT, maxT). prim: id or -1. t/uv/n may be NULL.
 * uv: barycentrics (b1, b2) for triangles. p: hit point, nrm: shading normal (closest hit only). */
int rtg_trace(rtg_ctx* ctx, rtg_scene* scene, const float* rays, int32_t n, int32_t any_hit, int32_t* prim,
              float* t, float* uv, float* p, float* nrm);
/* Allocation-free variant of rtg_trace. It has no counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc). rays_device (n x 8 floats) and the outputs are caller-owned DEVICE
 * buffers laid out as rtg_trace's (t, uv, p, nrm may be NULL); the call allocates nothing: the
 * traversal's scratch (its stack spill area and two counters, 18 MiB + 256 bytes) is a static
 * device array of the library, one per GPU, held until the process ends and shared by every call
 * on that GPU, so calls that overlap on one GPU (from different contexts) race. A scene whose
 * traversal stack bound (rtg_scene_info.traversal_stack) exceeds 48 entries fails with
 * RTG_ERR_INVALID. Returns when the results are written; rtg_trace_spills reports its spills. */
int rtg_trace_device_noalloc(rtg_ctx* ctx, rtg_scene* scene, const float* rays_device, int32_t n, int32_t any_hit,
                             int32_t* prim_device, float* t_device, float* uv_device, float* p_device,
                             float* nrm_device);
/* Traversal-stack pushes of the context's last rtg_trace call that went past the 12 on-chip (LDS)
 * entries of a ray's stack into the global spill area (no reference counterpart: the reference's
 * stack is a 64-entry array, accelStruct/Bvh.cpp:392-400). A stack deeper than the BVH builder's
 * bound fails the render or trace with RTG_ERR_HIP instead of losing geometry. */
int rtg_trace_spills(rtg_ctx* ctx, uint64_t* spilled_pushes);

/* Sampler tap for Sampler::getNextSample2D / getNextSample1D (sampler/Sampler.h:18-31;
 * StratifiedSampler.cpp:18-124, RandomSampler.cpp:28-45): the draws the renderer's sampler gives
 * each of n (pixel index y * width + x, sample index) pairs of the render `desc` (sampler, spp —
 * rounded up to a square for the stratified sampler —, dimension, seed): out2 (n x n2 x 2) the
 * first n2 2D draws, out1 (n x n1) the first n1 1D draws. The GPU sampler is counter-based, not
 * the reference's sequential PCG32 stream; with the stratified sampler the first `dimension` 2D
 * draws of a pixel's samples cover its sqrt(spp) x sqrt(spp) strata once each, its first
 * `dimension` 1D draws its spp strata once each, as the reference's shuffled strata do. */
int rtg_sampler_batch(rtg_ctx* ctx, const rtg_render_desc* desc, const uint32_t* pixel_sample, int32_t n,
                      int32_t n2, int32_t n1, float* out2, float* out1);

/* Parity tap for BSDF::eval/pdf/sample (bsdf/BSDF.h:70-74) on local directions.
 * op 0: eval+pdf of (wi, wo) -> f (n x 3), pdf (n).
 * op 1: sample from wi with u (n x 2) -> wo (n x 3), f = weight (n x 3), pdf (n), rel_eta (n).
 * op 2 (diffuse and rough conductor; used by guided sampling, no reference counterpart): for
 *   (wi, wo) the density with which op 1 produces wo -> pdf, the weight op 1 returns for wo -> f,
 *   the pdf op 1 reports for wo -> rel_eta.
 * op 3 (rough dielectric; used by guided renders, no reference counterpart): the exit proposals
 *   with u (n x 3) = (technique choice, ux, uy) -> wo, f = the MIS-weighted estimate, pdf = the
 *   pdf the sampler reports for wo, rel_eta = the MIS factor applied to the sampler's weight.
 * Smooth conductors and dielectrics: op 0 gives f = 0 and pdf = 0 (their eval and pdf in the
 * solid-angle measure), op 1 the mirror / refracted direction with pdf 1, fr or 1 - fr. */
int rtg_bsdf_batch(rtg_ctx* ctx, const rtg_material* material, int32_t op, const float* wi, const float* wo_or_u,
                   int32_t n, float* out_dir, float* f, float* pdf, float* rel_eta);

/* Parity tap for the lights of next-event estimation and of the MIS weight at an emitter hit:
 * Light::sample / Light::pdf (light/Light.h:56-64; mesh emitters light/MeshLight.cpp:32-64 with
 * geometricShape/Mesh.cpp:38-93, the environment light/EnvironmentLight.cpp:56-135),
 * EnvironmentLight::le (light/EnvironmentLight.cpp:137-166) and the one-light strategy's choice
 * (integrator/Integrator.cpp:176-190 with core/CDF.h). `light` indexes the scene's lights: the
 * emitting meshes in upload order, then the environment (rtg_scene_info.lights in all).
 * op 0 sample: in = n points (x 3), u = n x 2 draws -> out n x 14: unit direction to the light (3),
 *   distance, pdf (solid angle, without the selection probability), intensity (3), sampled point
 *   (3), its normal (3; zero for the environment).
 * op 1 pdf: in = n x 12 (from point, point on the light, its normal, direction) -> out n.
 * op 2 environment radiance: in = n directions (x 3) -> out n x 3 (`light` is ignored).
 * op 3 light choice: in = n draws -> out n x 3: light index, the draw rescaled for reuse, selection
 *   probability (`light` is ignored). */
int rtg_light_batch(rtg_ctx* ctx, rtg_scene* scene, int32_t op, int32_t light, const float* in, int32_t n,
                    const float* u, float* out);
/* Parity tap for NoriPerspective::getRay (camera/NoriPerspective.cpp:80-118) with the camera of a
 * render description (origin/target/up or camera_to_world, fov, size, clip planes): pxy = n film
 * positions (pixel + jitter, x 2) -> rays n x 8 (origin, direction, minT, maxT), as the
 * regeneration stage makes its camera rays. */
int rtg_camera_rays(rtg_ctx* ctx, const rtg_render_desc* desc, const float* pxy, int32_t n, float* rays);

/* Parity tap for the scene's camera medium (medium/Heterogeneous.cpp, volume/GridVolume.cpp,
 * medium/HGPhaseFunction.cpp); the scene must have one, a grid one for ops 0-2.
 * op 0 GridVolume::lookupSigmaT x scale: in = n points (x 3) -> out n.
 * op 1 Heterogeneous::sampleDistance (delta tracking) and op 2 Heterogeneous::transmittance (ratio
 *   tracking): in = n rays (x 8: origin, direction, minT, maxT), u = k given 1D draws per ray (a ray
 *   that needs more draws gets 0.5) -> out n x 3: (scattered 0/1, t, draws used) for op 1,
 *   (transmittance, draws used, 0) for op 2.
 * op 3 HGPhaseFunction::eval: in = n x 6 (wi, wo) -> out n.
 * op 4 HGPhaseFunction::sample: in = n x 3 (wi), u = n x 2 (k = 2) -> out n x 4 (wo, pdf). */
int rtg_medium_batch(rtg_ctx* ctx, rtg_scene* scene, int32_t op, const float* in, int32_t n, const float* u,
                     int32_t k, float* out);

/* Path-guiding network parity taps (no reference counterpart; the network is this library's own).
 * params: RTG_GUIDE_PARAMS FP32 values, per layer W[out][in] row-major then bias:
 *   W1 64x144, b1 64, W2 64x64, b2 64, W3 64x64, b3 64, W4 40x64, b4 40.
 * The device computes in bf16 with FP32 accumulation (weights and activations rounded to bf16).
 * rtg_guide_eval: inputs in6 (n x 6: position in [0,1]^3, unit normal) -> mixture (n x 40: per lobe
 *   weight, mean direction xyz, concentration).
 * rtg_guide_grad: samples (n x 10: position, normal, direction, weight) -> loss
 *   -(1/sum w) sum w log q(direction) and its gradient (RTG_GUIDE_PARAMS floats); n <= 65536.
 * rtg_guide_train: `steps` Adam steps (lr) over the same samples starting from params;
 *   params_out gets the result, losses (steps entries, may be NULL) the loss before each step. */
enum { RTG_GUIDE_PARAMS = 20200 };
int rtg_guide_eval(rtg_ctx* ctx, const float* params, const float* in6, int32_t n, float* mix_out);
int rtg_guide_grad(rtg_ctx* ctx, const float* params, const float* samples, int32_t n, float* loss, float* grad);
int rtg_guide_train(rtg_ctx* ctx, const float* params, const float* samples, int32_t n, int32_t steps, float lr,
                    float* params_out, float* losses);
/* The network of the context's last guided render (RTG_GUIDE_PARAMS floats, layout as above). */
int rtg_guide_params(rtg_ctx* ctx, float* params_out);
/* Directions from mixtures (n x 40, as rtg_guide_eval returns) with u (n x 3: lobe choice, two for
 * the lobe) -> dirs (n x 3) and the mixture density at each (n), through the layout and the sampling
 * and density functions guided renders use. */
int rtg_guide_sample(rtg_ctx* ctx, const float* mix, const float* u, int32_t n, float* dirs, float* pdf);

/* Continuous-batching serving (no reference counterpart; it replaces the reference's job loop,
 * JobManager::addJobs / jobRun in core/JobManager.cpp:103-232, for a host that renders many
 * requests on one GPU, e.g. an interactive preview service). Requests of any resident scene —
 * PathTracingMIS or VolPathTracing — are admitted in arrival order while a job slot (256) and the
 * HBM budget for their FP32 films allow; every wavefront iteration hands free path slots to the
 * admitted requests, shortest remaining work first, so their paths share the queues. A request's
 * image equals its rtg_render image for the same seed up to the order of the film's float
 * additions. Steady-state iteration batches are replayed hipGraphs.
 * While a server exists on a context, rtg_render / rtg_render_device on that context become
 * requests of the server and wait for their result; guided renders and all-lights PathTracingMIS
 * renders (which the server does not take) fail with RTG_ERR_INVALID.
 * hbm_budget_bytes: film bytes that may be resident at once (0 = 8 GiB); a request whose film
 * alone exceeds it is cancelled at admission. */
typedef struct rtg_server rtg_server;
int rtg_server_create(rtg_ctx* ctx, uint64_t hbm_budget_bytes, rtg_server** out);
/* Stops the server thread if it runs, frees every request's film. */
int rtg_server_destroy(rtg_server* srv);
/* Start the server's own loop thread: it runs the device loop whenever requests are pending or
 * running (slices of 20 ms under the context lock) and sleeps otherwise. Afterwards
 * rtg_server_step fails; use rtg_wait / rtg_poll. */
int rtg_server_start(rtg_server* srv);
/* Enqueue a request (validated now); returns its id in *job_id. Does not block. */
int rtg_submit(rtg_server* srv, rtg_scene* scene, const rtg_render_desc* desc, int64_t* job_id);
/* Run the device loop on the calling thread until no request is pending or running, or max_ms
 * elapsed (0 = until idle). Requests submitted from another thread between calls join at the next
 * iteration batch. */
int rtg_server_step(rtg_server* srv, double max_ms);
/* RTG_OK when finished (copies width*height*4 floats of (rgb sum, weight) if out_rgbw != NULL),
 * RTG_ERR_NOT_READY while queued or running, RTG_ERR_CANCELLED after a cancel. *latency_ms =
 * submit-to-finish wall time. */
int rtg_poll(rtg_server* srv, int64_t job_id, float* out_rgbw, double* latency_ms);
/* Block until the request has finished or was cancelled (results as rtg_poll), or timeout_ms
 * passed (RTG_ERR_NOT_READY; timeout_ms <= 0 waits without limit). Without a server thread the
 * calling thread runs the device loop meanwhile. */
int rtg_wait(rtg_server* srv, int64_t job_id, double timeout_ms, float* out_rgbw, double* latency_ms);
/* Stop claiming new samples for the request; its paths in flight drain, then it ends cancelled. */
int rtg_cancel(rtg_server* srv, int64_t job_id);

typedef struct rtg_server_stats {
  uint64_t submitted, admitted, completed, cancelled;
  uint64_t iterations;          /* wavefront iterations run by the loop */
  uint64_t samples;             /* camera samples of finished requests */
  uint64_t film_bytes;          /* film bytes resident now */
  uint64_t film_bytes_peak;     /* most film bytes resident at once */
  int32_t running;              /* requests admitted and not finished */
  int32_t pending;              /* requests waiting for admission */
  int32_t max_running;          /* most requests sharing one iteration */
  int32_t reserved;
  uint64_t graph_launches;      /* iteration batches replayed from hipGraphs on the context */
} rtg_server_stats;
int rtg_server_query(rtg_server* srv, rtg_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* RTG_H */
