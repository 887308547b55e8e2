/*
 * oracle.h -- CPU restatement of Asuna's render path.  TEST INFRASTRUCTURE
 * ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * call it, always as the checker.  The product (libasuna.so) never links,
 * loads or calls anything here.
 *
 * Every entry point restates a reference function op for op (citations are on
 * the definitions in oracle.cpp).  The estimator is templated on the real
 * type: precision 0 = FP32 (faithful to the GLSL, RNG bit-exact), precision 1
 * = FP64 (the same estimator in double, used as the converged reference).
 *
 * Parity pins (what checks each function, see tests/test_oracle_*.py):
 *   RNG (xxhash32Seed, pcg, rand)      hand-computed known answers + published
 *                                       xxHash32/PCG constants
 *   camera chain                        closed forms (centre ray = forward axis,
 *                                       edge rays at tan(fov/2), aspect)
 *   watertight intersection             brute force == BVH; analytic hits
 *   sampling warps                      chi-square against the stated pdf
 *   BxDFs                               white furnace, reciprocity, Fresnel
 *                                       special cases, GGX normalisation
 *   env tables                          constant env closed form; bilinear
 *                                       REPEAT texel blends by hand
 *   accumulation                        weighted-mean recurrence closed form
 *   end to end                          furnace/closed-box analytic radiance
 */
#ifndef ASUNA_ORACLE_H
#define ASUNA_ORACLE_H

#include <stdint.h>
#include "../include/asuna_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* RNG (math.glsl:20-44). */
uint32_t orc_xxhash32_seed(uint32_t x, uint32_t y, uint32_t z);
uint32_t orc_pcg(uint32_t* state);
void orc_rand_stream(uint32_t seed, int n, float* out);

/* Camera (camera.cpp:28-98 + nvmath look_at / invert / invert_rot_trans). */
void orc_camera_perspective(uint32_t w, uint32_t h, float fov_deg, const float eye[3],
                            const float lookat[3], const float up[3], float aperture,
                            float focal_distance, const float env_transform[16],
                            GpuCamera* out);
/* Primary ray of pixel (x,y) with jitter (jx,jy), no lens sample (rgen:30-90). */
void orc_camera_ray(const GpuCamera* cam, float px, float py, float o[3], float d[3]);
/* Camera ray of either camera type at raster (px,py); a perspective camera
 * with aperture > 0 takes the lens sample (l0,l1) (rgen:30-87). */
void orc_camera_ray_lens(const GpuCamera* cam, float px, float py, float l0, float l1, float o[3],
                         float d[3]);

/* Env tables (texture.cpp:144-226): marginal H x 2 (x,y of column 0),
 * conditional W*H x 2. */
void orc_env_tables(uint32_t w, uint32_t h, const float* rgba, float* marginal,
                    float* conditional);
/* Bilinear LOD-0 REPEAT texture fetch emulation of a 4-channel image. */
void orc_texture_bilinear(uint32_t w, uint32_t h, const float* rgba, float u, float v,
                          float out[4]);

/* Scene: flattened world triangles + binned-SAH BVH (the oracle's own). */
int orc_scene_create(const ash_scene_desc* desc, orc_scene** out);
void orc_scene_destroy(orc_scene* s);
uint64_t orc_scene_num_triangles(const orc_scene* s);

/* Closest hit / any hit of one ray (traceRayEXT contract, rgen:108-122). */
int orc_trace_closest(const orc_scene* s, const float o[3], const float d[3], float tmin,
                      float tmax, ash_hit* out, int brute_force);
int orc_trace_any(const orc_scene* s, const float o[3], const float d[3], float tmin,
                  float tmax, int brute_force);

/* Frame-0 primary hits of a w x h film (first-hit parity buffer): pixel
 * centres, plus the frame-0 lens draws of a thin-lens camera. */
int orc_first_hit(const orc_scene* s, const GpuCamera* cam, uint32_t w, uint32_t h,
                  ash_hit* out, int threads);

/* Progressive render of frames [first_frame, first_frame+n_frames) exactly as
 * runOffline does (tracer.cpp:218-226 + rgen:26-179), accumulating into
 * images (9 x w x h x 4 floats, read and written).  Only the pixels in rows
 * [y0,y1) and columns [x0,x1) are rendered (crops for the FP64 oracle).
 * precision 0 = FP32, 1 = FP64.  var_out (optional, w*h*4 floats) receives
 * per-pixel sums {sum w, sum w*L.r, sum w*L.g, sum w*L.b} is NOT provided;
 * instead sq_out (optional, w*h*4) accumulates {sum L.r^2, sum L.g^2,
 * sum L.b^2, count} of the clamped per-sample radiance, for variance
 * estimates.  stats (optional, 4 uint64) receives {paths, extension rays,
 * shadow rays, dielectric shadow rays}. */
int orc_render(const orc_scene* s, const GpuCamera* cam, const GpuPushConstantRaytrace* pc,
               uint32_t w, uint32_t h, int32_t first_frame, int32_t n_frames,
               uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, int precision,
               int threads, float* images, float* sq_out, uint64_t* stats,
               int32_t depth_channel);

/* Single-function probes for unit tests (FP32 path). */
void orc_concentric_disk(float u0, float u1, float out[2]);
void orc_cosine_hemisphere(float u0, float u1, float out[3]);
void orc_offset_ray(const float p[3], const float n[3], float out[3]);
void orc_basis(const float n[3], float f[3], float r[3]);
float orc_power_heuristic(float a, float b);
float orc_compute_diffuse_fresnel(float ior, int sample_count);
/* BxDF probes: material type, local-frame inputs (N = +z, X = +x, Y = +y). */
void orc_bsdf_eval(const GpuMaterial* m, const float L[3], const float V[3], uint32_t flags,
                   float out[3]);
float orc_bsdf_pdf(const GpuMaterial* m, const float L[3], const float V[3], uint32_t flags);
/* The rough-plastic light-sample terms (brdf_rough_plastic.rchit:229-241):
 * eval and pdf with the shader's permuted arguments, as the render uses them. */
void orc_rough_plastic_light_terms(const GpuMaterial* m, const float L[3], const float V[3], uint32_t light_flags,
                                   float bw[3], float* bsdf_pdf);
/* Shirley's conductor Fresnel (brdf_conductor.rchit:14-34). */
float orc_conductor_reflectance(float eta, float k, float cos_theta_i);
/* Sun & sky (utils/sun_and_sky.glsl): set the scene's record; evaluate the
 * sky radiance toward dir (FP32, or FP64 when fp64 != 0); the physical sun
 * scales (disk, glow) of calc_physical_scale. */
void orc_set_sun_and_sky(orc_scene* s, const GpuSunAndSky* sky);
void orc_sun_and_sky(const GpuSunAndSky* sky, const float dir[3], int fp64, float out[3]);
void orc_sky_physical_scale(float disk_scale, float glow_intensity, float disk_intensity, float out[2]);
/* Sample: draws from *seed in reference order; returns weight (f*cos, or
 * f*cos/pdf where the shader divides later is NOT applied), direction,
 * pdf, flags. */
void orc_bsdf_sample(const GpuMaterial* m, const float V[3], uint32_t* seed, float weight[3],
                     float dir[3], float* pdf, uint32_t* flags);
void orc_sample_envmap(const orc_scene* s, const float env_transform[16], float intensity,
                       float r0, float r1, float L[3], float* pdf, float rad[3]);
float orc_pdf_envmap(const orc_scene* s, const float env_transform[16], const float L[3]);
void orc_eval_envmap(const orc_scene* s, const float env_transform[16], float intensity,
                     const float L[3], float rad[3]);

#ifdef __cplusplus
}
#endif
#endif
