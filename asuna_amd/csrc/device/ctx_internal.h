// Library-internal access to a render context (used by the render server).
#pragma once

#include "render.h"
#include "../../../include/asuna_denoise.h"

struct ash_ctx;

namespace ashd {

hipStream_t ctxStream(ash_ctx* c);
int ctxNumCU(ash_ctx* c);
// Wavefront buffers for `capacity` paths and traversal stacks for `grid` blocks.
int ctxPrepare(ash_ctx* c, uint32_t capacity, uint32_t grid);
int ctxCheckPc(ash_ctx* c, const GpuPushConstantRaytrace* pc);
// Ray generation, bounce loop and accumulation of one batch on the context's stream.
int ctxRunBatch(ash_ctx* c, const BatchDesc& b, int maxDepth, uint32_t grid, const ash_render_opts* o,
                ash_stats* st);
ShotParams ctxMakeShot(const GpuCamera* cam, const GpuPushConstantRaytrace* pc, int32_t depthChannel,
                       float4* images, uint32_t w, uint32_t h);
const char* ctxError(ash_ctx* c);
// Film size of the single-shot output images (0 x 0 before ash_set_film).
void ctxFilm(ash_ctx* c, uint32_t* w, uint32_t* h);

// ash_denoise_planes without the wait: enqueues the denoise on the
// denoiser's stream between the caller's events `start` and `done`.
int denoisePlanesAsync(ash_denoiser* d, const void* radiance, const void* albedo, const void* normal,
                       const void* depth, uint32_t W, uint32_t H, void* out, hipEvent_t start, hipEvent_t done);

}  // namespace ashd
