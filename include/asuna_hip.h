/*
 * asuna_hip.h -- C ABI of the MI355X render library (libasuna.so).
 *
 * This is the seam that replaces Asuna's Vulkan ray-tracing pipeline:
 *   PipelineRaytrace::init  (src/pipeline/pipeline_raytrace.cpp:7-23)
 *   PipelineRaytrace::run   (src/pipeline/pipeline_raytrace.cpp:36-74)
 *   PipelineGraphics::run / updateCameraBuffer (src/pipeline/pipeline_graphics.cpp:29-98)
 *   Scene::submit device uploads (src/scene/scene.cpp:26-85)
 * together with the POD structs those read (the headers under src/shared).  Every struct
 * below is byte-identical to the GLSL/C++ shared header it names; sizes and
 * offsets are checked by static assertions in the library and by the CPU
 * test-suite through ash_pod_layout().
 *
 * Conventions (differences from the reference are deliberate):
 *  - every function returns ASH_OK (0) or a negative ASH_E* code and never
 *    exits the process (the reference calls exit(1), e.g. scene.cpp:320-348);
 *    the message of the last failure is ash_last_error(ctx).
 *  - the context owns all device memory; callers own host inputs, which are
 *    copied during the call.
 *  - there is no process-global camera (CameraManip, camera.h:27): the camera
 *    and the push-constant state are passed to every render call.
 */
#ifndef ASUNA_HIP_H
#define ASUNA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ PODs */

/* GpuVertex, src/shared/vertex.h:6-11 (44 B, scalar layout). */
typedef struct GpuVertex {
  float pos[3];
  float uv[2];
  float normal[3];
  float tangent[3];
} GpuVertex;

/* GpuMaterial, src/shared/material.h:24-49 (132 B). */
typedef struct GpuMaterial {
  float diffuse[3];
  float rhoSpec[3];
  float anisoAlpha[2];
  float ior;
  float roughness;
  float subsurface;
  float specular;
  float specularTint;
  float anisotropic;
  float sheen;
  float sheenTint;
  float clearcoat;
  float clearcoatGloss;
  float radiance[3];
  float metalness;
  float radianceFactor[3];
  int32_t diffuseTextureId;
  int32_t roughnessTextureId;
  int32_t metalnessTextureId;
  int32_t radianceTextureId;
  int32_t normalTextureId;
  int32_t tangentTextureId;
  int32_t opacityTextureId;
  uint32_t type;
} GpuMaterial;

/* MaterialType, src/shared/material.h:7-21. */
enum {
  MaterialTypeBrdfLambertian = 0,
  MaterialTypeBrdfKang18 = 1,
  MaterialTypeBrdfEmissive = 2,
  MaterialTypeBrdfPbrMetalnessRoughness = 3,
  MaterialTypeBrdfPlastic = 4,
  MaterialTypeBrdfRoughPlastic = 5,
  MaterialTypeBrdfConductor = 6,
  MaterialTypeBrdfRoughConductor = 7,
  MaterialTypeBrdfMirror = 8,
  MaterialTypeBrdfDisney = 9,
  MaterialTypeBsdfDielectric = 10,
  MaterialTypeBrdfPhong = 11,
  MaterialTypeNum = 12
};

/* GpuLight, src/shared/light.h:16-26 (76 B). */
typedef struct GpuLight {
  int32_t type;
  float position[3];
  float direction[3];
  float radiance[3];
  float u[3];
  float v[3];
  float radius;
  float area;
  uint32_t doubleSide;
} GpuLight;

/* LightType, src/shared/light.h:7-13. */
enum {
  LightTypeDirectional = 0,
  LightTypeRect = 1,
  LightTypeTriangle = 2,
  LightTypePoint = 3,
  LightTypeUndefined = 4
};

/* GpuCamera, src/shared/camera.h:15-24 (224 B).  Matrices are column-major
 * (nvmath mat4f / GLSL mat4 storage: element (row r, col c) at [c*4 + r]). */
typedef struct GpuCamera {
  float rasterToCamera[16];
  float cameraToWorld[16];
  float envTransform[16];
  float fxfycxcy[4];
  uint32_t type;
  float aperture;
  float focalDistance;
  float padding;
} GpuCamera;

/* CameraType, src/shared/camera.h:7-11. */
enum { CameraTypePerspective = 0, CameraTypeOpencv = 1, CameraTypeUndefined = 2 };

/* GpuPushConstantRaytrace, src/shared/pushconstant.h:10-34 (84 B). */
typedef struct GpuPushConstantRaytrace {
  int32_t spp;
  int32_t curFrame;
  int32_t maxPathDepth;
  int32_t numLights;
  float bgColor[3];
  uint32_t useFaceNormal;
  uint32_t ignoreEmissive;
  uint32_t hasEnvMap;
  float envMapResolution[2];
  float envMapIntensity;
  uint32_t nMultiChannel;
  int32_t diffuseOutChannel;
  int32_t specularOutChannel;
  int32_t roughnessOutChannel;
  int32_t normalOutChannel;
  int32_t positionOutChannel;
  int32_t tangentOutChannel;
  int32_t uvOutChannel;
} GpuPushConstantRaytrace;

/* GpuSunAndSky, src/shared/sun_and_sky.h:6-25 (96 B): the procedural sky
 * (NVIDIA's MDL sun & sky, utils/sun_and_sky.glsl).  When in_use == 1 it
 * replaces the env map and background in the miss shader
 * (raytrace.default.rmiss:29-46) and is sampled uniformly over the sphere in
 * NEE (rchit_layouts.glsl:98-116, sample_light.glsl:137-141). */
typedef struct GpuSunAndSky {
  float rgb_unit_conversion[3];
  float multiplier;
  float haze;
  float redblueshift;
  float saturation;
  float horizon_height;
  float ground_color[3];
  float horizon_blur;
  float night_color[3];
  float sun_disk_intensity;
  float sun_direction[3];
  float sun_disk_scale;
  float sun_glow_intensity;
  int32_t y_is_up;
  int32_t physically_scaled_sun;
  int32_t in_use;
} GpuSunAndSky;

/* Number of RGBA32F output images, src/shared/binding.h:70.
 * image 0 = running weighted mean of radiance, images 1..7 = multi-channel
 * AOVs (channel c -> image c+1), image 8 = sum of filter weights
 * (raytrace.projective.rgen:155-178).  Channel 7 aliases image 8 exactly as
 * in the reference. */
#define ASH_NUM_OUTPUT_IMAGES 9

/* Extension channel (not in the reference): first-hit depth (ray t) written
 * like an AOV when the push constant's extension slot asks for it; see
 * ash_render_opts.depthOutChannel. */

/* --------------------------------------------------------- scene upload */

/* One mesh as Mesh/MeshAlloc hold it (src/core/mesh.cpp:51-101): de-indexed
 * GpuVertex array plus a uint32 index triple per triangle. */
typedef struct ash_mesh_desc {
  uint32_t num_vertices;
  uint32_t num_indices; /* multiple of 3 */
  const GpuVertex* vertices;
  const uint32_t* indices;
} ash_mesh_desc;

/* One TLAS instance (src/core/instance.h:10-31, pipeline_raytrace.cpp:120-147).
 * Instances with light_id >= 0 are emitters and always shade with the
 * Lambertian light-hit path (pipeline_raytrace.cpp:137-140). */
typedef struct ash_instance_desc {
  float transform[16]; /* object->world, column-major; row 3 must be 0,0,0,1 */
  uint32_t mesh_id;
  uint32_t material_id;
  int32_t light_id;
  uint32_t pad;
} ash_instance_desc;

/* RGBA32F image, row-major, row 0 first (readImage, src/core/texture.cpp:306-338). */
typedef struct ash_image_desc {
  uint32_t width;
  uint32_t height;
  const float* rgba;
} ash_image_desc;

/* Everything Scene::submit uploads (scene.cpp:41-72). Light 0 and material 0
 * are the dummies Scene::reset adds (scene.cpp:93-111). */
typedef struct ash_scene_desc {
  uint32_t num_meshes;
  const ash_mesh_desc* meshes;
  uint32_t num_instances;
  const ash_instance_desc* instances;
  uint32_t num_materials;
  const GpuMaterial* materials;
  uint32_t num_lights;
  const GpuLight* lights;
  uint32_t num_textures;
  const ash_image_desc* textures; /* texture id -> image */
  ash_image_desc envmap;          /* rgba==NULL: no env map (1x1 dummy) */
} ash_scene_desc;

/* --------------------------------------------------------------- results */

/* Frame-0 primary hit of one pixel (first-hit parity buffer). */
typedef struct ash_hit {
  float t;          /* ray parameter, INFINITY(1e10) on miss */
  float u, v;       /* barycentrics of vertex 1 and 2 (GLSL _bary) */
  int32_t inst_id;  /* gl_InstanceID, -1 on miss */
  int32_t prim_id;  /* gl_PrimitiveID within the instance's mesh, -1 on miss */
  float normal[3];  /* shading normal N of getHitState (rchit_layouts.glsl:67-95) */
  float pos[3];     /* world hit position of getHitState */
} ash_hit;

typedef struct ash_stats {
  uint64_t paths;          /* camera samples traced (pixels x frames x shots) */
  uint64_t extension_rays; /* closest-hit rays traced */
  uint64_t shadow_rays;    /* any-hit rays traced */
  uint64_t dielectric_shadow_rays; /* of which from delta lobes (dielectric, mirror, conductor):
                                      always zero contribution, counted but not traced */
  double render_ms;        /* device time of the render call (HIP events) */
  double extend_ms;        /* sum of closest-hit kernel time, if timing enabled */
  double shadow_ms;        /* sum of any-hit kernel time, if timing enabled */
  double shade_ms;         /* sum of shading kernel time, if timing enabled */
  uint64_t extend_launches; /* closest-hit kernel launches */
  uint64_t shadow_launches; /* any-hit kernel launches */
  /* traversal work, filled when ash_render_opts.count_traversal = 1 */
  uint64_t ext_node_visits; /* BVH nodes fetched by closest-hit rays */
  uint64_t ext_tri_tests;   /* triangles fetched and tested by closest-hit rays */
  uint64_t shadow_node_visits;
  uint64_t shadow_tri_tests;
  uint64_t batches;        /* wavefront batches */
  uint64_t graph_launches; /* bounce loops replayed from a captured HIP graph */
  /* Fused trace launches (default traversal scheme): one launch traces the
   * shadow rays of a bounce and the extension rays of the next; the first
   * bounce's extension rays and the last bounce's shadow rays keep their own
   * launches (counted in extend_launches / shadow_launches). */
  uint64_t trace_launches;
  double trace_ms;         /* their kernel time, if timing enabled */
  uint64_t trace_ext_rays, trace_shadow_rays; /* rays they traced, when count_traversal = 1 */
  uint64_t trace_ext_node_visits, trace_ext_tri_tests, trace_shadow_node_visits, trace_shadow_tri_tests;
} ash_stats;

typedef struct ash_build_opts {
  uint32_t leaf_max_tris; /* max triangles per leaf after collapse (default 4) */
  uint32_t quality;       /* 0 = LBVH, 1 = LBVH + SAH leaf collapse, 2 = PLOC (locally-ordered clustering
                             over the Morton order, radius 16) + SAH leaf collapse; a null opts pointer
                             selects 2 */
  uint32_t branching;     /* 2 = binary BVH, 4 = 4-wide BVH traversed by the kernels (default 4) */
  uint32_t node_format;   /* 4-wide nodes the wave-fetch kernels walk: 0 = default (float),
                             1 = 128-byte float boxes, 2 = 64-byte quantized boxes (8-bit planes on a
                             per-node power-of-two grid, rounded outwards: the same closest hits) */
  uint32_t two_level;     /* 1 = instancing without duplication: one object-space 4-wide BVH per distinct
                             mesh (instances sharing vertex and index ranges share it) under a binary tree
                             over the instances' world boxes (needs branching 4 and float nodes); the
                             flat default is one world-space tree over every instance's triangles */
} ash_build_opts;

typedef struct ash_build_info {
  uint64_t num_triangles;
  uint64_t num_nodes;
  uint32_t max_depth;
  uint32_t node_bytes; /* bytes of one node of the tree the kernels walk (64 binary, 128 or 64 4-wide) */
  double build_ms;
  double sah_cost;
  uint64_t device_bytes; /* nodes + triangles + shading data */
  uint64_t num_wide_nodes; /* 4-wide nodes (0 when branching = 2) */
  uint32_t wide_depth;     /* levels of the 4-wide tree */
  uint32_t branching;      /* tree the kernels traverse */
  uint32_t blas_count;     /* two-level: object-space trees built (0 when flat) */
  uint64_t blas_triangles; /* two-level: triangles stored in them (distinct meshes, not instances) */
} ash_build_info;

/* Render options beyond the push constants. */
typedef struct ash_render_opts {
  uint32_t queue_paths;    /* paths per wavefront batch, the tile size of the work split (0 = 2^25; the
                              wavefront is sized by the largest batch a render needs) */
  uint32_t blocks_per_cu;  /* persistent traversal blocks (256 threads) per CU (0 = 8) */
  uint32_t timing;         /* 1 = time each stage with HIP events */
  int32_t depthOutChannel; /* extension: AOV channel receiving first-hit t, -1 = none */
  uint32_t max_bounce_iters; /* safety cap on wavefront iterations (0 = auto) */
  uint32_t count_traversal;  /* 1 = count node visits / triangle tests (slower) */
  uint32_t no_graphs;        /* 1 = launch the bounce loop eagerly instead of replaying a HIP graph */
  uint32_t lds_stack;        /* closest-hit stack entries per thread held in LDS: 8 or 12 (0 = 12) */
  uint32_t traversal;        /* ray fetch of the traversal kernels: 0 = per-XCD slices of the queue
                                (default; closest hit on float 4-wide nodes with dynamic per-lane fetch
                                and speculative while-while rounds), 1 = wave fetch in queue order,
                                2 = per-lane refill, one node or triangle per step (4-wide tree) */
} ash_render_opts;

/* ------------------------------------------------------------ error codes */
enum {
  ASH_OK = 0,
  ASH_E_INVALID = -1,    /* bad argument */
  ASH_E_HIP = -2,        /* HIP runtime failure */
  ASH_E_STATE = -3,      /* call out of order (e.g. render before upload) */
  ASH_E_NOMEM = -4,      /* device allocation failed */
  ASH_E_IO = -5,         /* file read / write failure */
  ASH_E_PARSE = -6,      /* malformed scene file */
  ASH_E_UNSUPPORTED = -7 /* feature outside the supported set */
};

typedef struct ash_ctx ash_ctx;

/* Context on one device (--gpu_id, main.cpp:20 -> context.cpp:304). */
int ash_create(int device_ordinal, ash_ctx** out);
void ash_destroy(ash_ctx* ctx);
const char* ash_last_error(const ash_ctx* ctx);

/* Upload a scene (Scene::submit, scene.cpp:26-85).  Replaces any previous one.
 * Env-map importance tables are built here (EnvMap::EnvMap, texture.cpp:144-226). */
int ash_upload_scene(ash_ctx* ctx, const ash_scene_desc* scene);

/* The defaults Scene::init writes (scene.cpp:112-129; in_use = 0). */
void ash_sun_and_sky_defaults(GpuSunAndSky* out);

/* Set the sun & sky record of the uploaded scene (replaces the GUI's edit of
 * Scene::getSunsky and its uniform upload, scene.cpp:437,
 * pipeline_graphics.cpp:100-103).  ash_upload_scene resets it to the
 * defaults, as each Scene starts from them.  Takes effect for later renders,
 * including ones replayed from cached graphs. */
int ash_set_sun_and_sky(ash_ctx* ctx, const GpuSunAndSky* sky);

/* Build the acceleration structure on the device (replaces
 * createBottomLevelAS/createTopLevelAS, pipeline_raytrace.cpp:107-147). */
int ash_build_accel(ash_ctx* ctx, const ash_build_opts* opts, ash_build_info* info);

/* Allocate/clear the 9 output images for a film of w x h
 * (PipelineGraphics::createOffscreenResources, pipeline_graphics.cpp:105-196). */
int ash_set_film(ash_ctx* ctx, uint32_t width, uint32_t height);

/* Render frames [first_frame, first_frame + n_frames) of one shot
 * (n_frames calls of PipelineRaytrace::run with curFrame = first_frame + i,
 * tracer.cpp:218-226), accumulating into the output images exactly as the
 * reference's running weighted mean.  pc->curFrame and pc->spp are ignored
 * (the offline loop forces spp = 1, tracer.cpp:211). */
int ash_render(ash_ctx* ctx, const GpuCamera* cam, const GpuPushConstantRaytrace* pc,
               int32_t first_frame, int32_t n_frames, const ash_render_opts* opts,
               ash_stats* stats);

/* Allocation-free variant of ash_render.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It renders frames [first_frame, first_frame + n_frames) of
 * one shot into the caller's device buffer `images`: ASH_NUM_OUTPUT_IMAGES
 * planes of width*height RGBA32F, laid out like ash_render's film and
 * accumulated the same way (frame 0 overwrites; later frames fold into the
 * running mean, so a caller continuing a render passes the same buffer).
 * The wavefront lives in one fixed 128 MiB device array per GPU, held
 * until the process ends; at most 2^18 paths are in flight per batch and
 * the traversal stacks of 2 persistent blocks per CU must fit beside them,
 * else ASH_E_INVALID.  No hipMalloc / hipFree is called.  Every call on a
 * GPU shares that GPU's array and keeps nothing from the previous call;
 * calls on one GPU are serialised (a second caller waits).  Results are
 * bit-identical to ash_render's.  Returns after the render completes;
 * stats (optional) carry counts and host-timed render_ms. */
int ash_render_noalloc(ash_ctx* ctx, const GpuCamera* cam, const GpuPushConstantRaytrace* pc,
                       int32_t first_frame, int32_t n_frames, uint32_t width, uint32_t height,
                       void* images, ash_stats* stats);

/* Batched multi-shot render: n_shots shots of one scene, each with its own
 * camera and push constants, all frames [0, n_frames) in one wavefront.
 * Output images of shot s are kept on the device; read them with
 * ash_read_shot_output. */
int ash_render_shots(ash_ctx* ctx, uint32_t n_shots, const GpuCamera* cams,
                     const GpuPushConstantRaytrace* pcs, int32_t n_frames,
                     const ash_render_opts* opts, ash_stats* stats);
int ash_read_shot_output(ash_ctx* ctx, uint32_t shot, int32_t image, float* host_rgba);

/* Autotuner (no counterpart in the reference, whose dispatch shape is fixed
 * at one thread per pixel): on a probe render of `probe_frames` frames of
 * the given shots (each candidate rendered twice, the second timed) it
 * sweeps the wavefront batch size queue_paths over 2^18, 2^19, ... up to
 * the first size at which the probe fits in one batch (at most 2^26), the
 * persistent traversal grid blocks_per_cu over {4, 8, 16} and the LDS share
 * of the closest-hit stack over {8, 12} with the traversal fetch scheme of
 * `inout`, then tries the other fetch schemes at the best of those, and
 * writes the fastest setting into `inout`.  The probe renders overwrite the
 * shot outputs. */
typedef struct ash_tune_result {
  uint32_t queue_paths;
  uint32_t blocks_per_cu;
  uint32_t candidates;
  uint32_t lds_stack;
  uint32_t traversal;
  uint32_t pad;
  double best_ms;  /* probe render time of the chosen setting */
  double worst_ms; /* slowest candidate */
} ash_tune_result;
int ash_autotune(ash_ctx* ctx, uint32_t n_shots, const GpuCamera* cams, const GpuPushConstantRaytrace* pcs,
                 int32_t probe_frames, ash_render_opts* inout, ash_tune_result* result);
/* Written into `inout`: queue_paths, blocks_per_cu, lds_stack, traversal. */

/* Copy output image `image` (0..8) to host RGBA32F (w*h*4 floats). */
int ash_read_output(ash_ctx* ctx, int32_t image, float* host_rgba);

/* Device pointer of output image `image` (RGBA32F, w*h*4 floats), for callers
 * that keep the result in HBM (denoiser, serving). */
int ash_output_device_ptr(ash_ctx* ctx, int32_t image, void** dev_ptr);

/* Frame-0 primary hits for every pixel: raygen with jitter 0.5 (rgen:41) +
 * one closest-hit trace + hit-state reconstruction.  out has w*h entries. */
int ash_first_hit(ash_ctx* ctx, const GpuCamera* cam, ash_hit* out);

/* Environment importance tables as the device uses them: marginal (H x 2
 * floats: row/H, pdf1D) and conditional (W*H x 2 floats: col/W, pdf2D). */
int ash_env_tables(ash_ctx* ctx, float* marginal_out, float* conditional_out);

/* Size / offset table of every POD, for layout checks:
 * writes pairs {sizeof, offsetof(last field)} for GpuVertex, GpuMaterial,
 * GpuLight, GpuCamera, GpuPushConstantRaytrace (10 uint32). */
void ash_pod_layout(uint32_t* out10);

/* Library build string (arch, flags). */
const char* ash_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ASUNA_HIP_H */
