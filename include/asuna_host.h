/*
 * asuna_host.h -- C ABI of the host side of libasuna.so: scene-file loading
 * (the reference's Loader + Scene, src/loader and src/scene), camera/shot
 * state (src/core/camera.cpp, src/scene/scene.cpp:439-453), the offline
 * render loop (Tracer::runOffline, src/tracer/tracer.cpp:177-291) and the
 * image writers (src/core/texture.cpp:340-372).
 *
 * The render seam itself is in asuna_hip.h; this header is what the CLI
 * (asuna --scene --offline --out ...) and the Python harness drive.
 */
#ifndef ASUNA_HOST_H
#define ASUNA_HOST_H

#include "asuna_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ash_host_scene ash_host_scene;

/* Global (non-push-constant) parts of the reference State
 * (src/core/state.h:6-63, loader.cpp:148-234). */
typedef struct ash_scene_state {
  uint32_t film_width, film_height;
  uint32_t num_shots;
  uint32_t tone_mapping;        /* ToneMappingType (pushconstant.h:41-50) */
  uint32_t output_hdr;          /* state.output_hdr */
  uint32_t output_render_result;/* state.output_render_result */
  uint32_t channel_output_ldr;  /* bit c = channel c written as LDR png */
  int32_t depth_out_channel;    /* extension channel "depth", -1 if absent */
  uint64_t num_triangles;       /* over all instances */
} ash_scene_state;

/* Loader::loadSceneFromJson (loader.cpp:67-144) + Scene::submit's host part
 * (scene.cpp:26-85).  Relative asset paths resolve against the scene file's
 * directory (nvh::findFile with the scene dir).  On failure returns a
 * negative code and writes a message into err. */
int ash_scene_load(const char* json_path, ash_host_scene** out, char* err, size_t err_len);
void ash_scene_free(ash_host_scene* s);

/* Views into the loaded scene, valid until ash_scene_free. */
int ash_scene_get_desc(const ash_host_scene* s, ash_scene_desc* out);
int ash_scene_get_state(const ash_host_scene* s, ash_scene_state* out);

/* Scene::setShot + CameraPerspective/CameraOpencv::toGpuStruct
 * (scene.cpp:439-453, camera.cpp:64-98): camera UBO and push constants of
 * shot `shot`, with curFrame = 0 and spp = 1 as the offline loop sets them. */
int ash_scene_shot(const ash_host_scene* s, uint32_t shot, GpuCamera* cam,
                   GpuPushConstantRaytrace* pc);

/* Camera helper: perspective GpuCamera for a look-at pose
 * (Camera::setToWorld + CameraPerspective::toGpuStruct). */
void ash_camera_lookat(uint32_t width, uint32_t height, float fov_deg, const float eye[3],
                       const float lookat[3], const float up[3], float aperture,
                       float focal_distance, const float env_transform16[16], GpuCamera* out);

/* Tracer::runOffline: every shot, all spp frames, outputs written with the
 * reference naming <out>_shot_%04d.{exr,png} and
 * <out>_shot_%04d_channel_%04d.{exr,png} (tracer.cpp:266-291); with
 * output_scanline the channel images are .npy of valid pixels
 * (tracer.cpp:366-392).  spp_override > 0 replaces every shot's spp (the
 * reference ignores --spp, scene.cpp:373; this is a documented divergence
 * that is off when spp_override <= 0). */
typedef struct ash_offline_opts {
  int32_t gpu_id;
  int32_t output_scanline;
  int32_t spp_override;
  int32_t quiet;
  ash_render_opts render;
} ash_offline_opts;
int ash_run_offline(const char* scene_json, const char* out_name, const ash_offline_opts* opts,
                    ash_stats* stats, char* err, size_t err_len);

/* writeImage (texture.cpp:340-372): .exr (half RGB), .hdr (RGBE), .png
 * (8-bit, x*255), .npy (float32 HxWx4). */
int ash_write_image(const char* path, uint32_t width, uint32_t height, const float* rgba);
/* readImage (texture.cpp:306-338): .hdr, .exr, .png; RGBA32F, caller frees
 * with ash_free_image.  gamma applies to 8-bit formats (stbi_ldr_to_hdr_gamma). */
int ash_read_image(const char* path, float gamma, uint32_t* width, uint32_t* height, float** rgba);
void ash_free_image(float* rgba);

/* EnvMap::EnvMap importance tables (texture.cpp:144-226) as ash_upload_scene
 * builds them: marginal H x 2 (column 0: row/H, pdf1D), conditional W*H x 2
 * (col/W, pdf2D). */
void ash_env_tables_host(uint32_t width, uint32_t height, const float* rgba, float* marginal,
                         float* conditional);

/* post.idle.frag tone mapping of a width x height RGBA32F image in place
 * (types none, gamma, reinhard, aces, filmic, pbrt, custom; custom with the
 * loader's default post state: no auto exposure). */
void ash_tonemap(uint32_t tm_type, uint32_t width, uint32_t height, float* rgba);

/* Procedural benchmark scenes written as JSON + OBJ (+ env map) into dir:
 *   config 0: closed box, 5 walls + 2 boxes + 0.25 x 0.25 rect light, 512^2,
 *             64 spp, depth 8 (lambertian).
 *   config 1: 8x8 grid of level-`level` icospheres displaced by 6-octave fBm,
 *             ground quad, env map 2048x1024 (sky + sun disk) + 4 triangle
 *             lights, pbr / dielectric materials, 1920x1080, 16 spp, depth 6.
 *   config 2: config 0's box seen from 64 shots (multi-shot batch), 32 spp,
 *             depth 6.
 * `seed` is the master seed (default 0xA5A52026); `scale` shrinks config 1
 * for tests (icosphere level and grid size, 0 = full size).  Writes the
 * scene file path into path_out. */
int ash_generate_scene(int32_t config, uint64_t seed, int32_t scale, const char* dir,
                       char* path_out, size_t path_len);

#ifdef __cplusplus
}
#endif
#endif
