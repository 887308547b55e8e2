/* Learned denoiser of libasuna.so.
 *
 * Stands where the reference's denoise stage stands (denoiser.cpp and its
 * call in tracer.cpp:560-562, which hand radiance + albedo + normal guides to
 * a closed library).  Here it is a 5-level U-Net of 3x3 convolutions (widths
 * 32/48/64/96/128, ReLU, 2x2 average pooling, nearest 2x upsampling, skip
 * concatenation) trained and run by the library's own BF16 MFMA kernels
 * (FP32 accumulation, FP32 master weights, Adam).
 *
 * Input features per pixel (10): log(1 + radiance) RGB, albedo RGB, normal
 * XYZ, log(1 + depth).  Output: radiance RGB (the network predicts it in
 * log(1 + x) space).  Loss: mean absolute error in that space.
 * Spatial sizes must be multiples of 16 for training; ash_denoise_film pads.
 */
#ifndef ASUNA_DENOISE_H
#define ASUNA_DENOISE_H

#include "asuna_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ash_denoiser ash_denoiser;

/* seed: parameter initialisation (He normal, zero bias). */
int ash_denoiser_create(int device, uint32_t seed, ash_denoiser** out);
void ash_denoiser_destroy(ash_denoiser* d);
const char* ash_denoiser_last_error(ash_denoiser* d);

/* Parameters as one FP32 vector: per convolution in network order,
 * W[tap][cout][cin] (tap = 3 * (dy + 1) + (dx + 1)) then bias[cout], real
 * channel counts.  ash_denoiser_set_params also resets the Adam moments. */
int ash_denoiser_num_params(ash_denoiser* d, uint64_t* n);
int ash_denoiser_get_params(ash_denoiser* d, float* host);
int ash_denoiser_set_params(ash_denoiser* d, const float* host);

/* Host-side batch: raw (N, H, W, 10) FP32 = radiance RGB, albedo RGB,
 * normal XYZ, depth; target (N, H, W, 3) FP32 radiance. */
int ash_denoiser_forward(ash_denoiser* d, uint32_t n, uint32_t h, uint32_t w, const float* input, float* output_rgb);
/* Loss and the FP32 gradient (same order as the parameters); no update. */
int ash_denoiser_gradient(ash_denoiser* d, uint32_t n, uint32_t h, uint32_t w, const float* input, const float* target,
                          float* loss, float* grad);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the batch; returns the loss before the step. */
int ash_denoiser_train_step(ash_denoiser* d, uint32_t n, uint32_t h, uint32_t w, const float* input,
                            const float* target, float lr, float* loss);

/* Denoise a render context's film in HBM: radiance = image 0, albedo /
 * normal / depth = AOV channels (image index = channel + 1; depth uses .x),
 * result written to image `out_image` (RGBA32F, alpha 1).  Runs on the
 * denoiser's stream after the context's pending work; returns after
 * completion.  ms = device time. */
int ash_denoise_film(ash_denoiser* d, ash_ctx* ctx, int32_t albedo_channel, int32_t normal_channel,
                     int32_t depth_channel, int32_t out_image, double* ms);

/* Same on raw device planes (RGBA32F, width x height; albedo / normal /
 * depth may be null = zero features).  The caller orders the planes'
 * producers before this call (it runs on the denoiser's own stream). */
int ash_denoise_planes(ash_denoiser* d, const void* radiance, const void* albedo, const void* normal,
                       const void* depth, uint32_t width, uint32_t height, void* out, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* ASUNA_DENOISE_H */
