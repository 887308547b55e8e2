/* Resident render server on one GPU: libasuna.so.
 *
 * No counterpart in the reference, which renders one scene per process and
 * one shot at a time (tracer.cpp:177-264).  Callers: render farms and
 * interactive services that keep several scenes resident and submit many
 * small jobs.
 *
 * Scenes stay in HBM (geometry, BVH, textures, env tables) once added.  Jobs
 * (camera, push constants, resolution, frames) are admitted when their
 * output images fit in the HBM budget, then co-batched: every wavefront
 * batch of a scene takes work units (pixel runs x frame ranges) from all of
 * that scene's admitted jobs, and the batches of different scenes run on
 * their own streams at the same time.  A job's result does not depend on
 * what it was batched with: the RNG is keyed by (pixel, frame) and every
 * pixel folds its frames in order, so it equals ash_render of the same
 * camera bit for bit.
 */
#ifndef ASUNA_SERVER_H
#define ASUNA_SERVER_H

#include "asuna_denoise.h"
#include "asuna_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ash_server ash_server;

typedef struct ash_server_opts {
  uint64_t hbm_budget_bytes; /* arena for admitted jobs' output images (0 = 8 GiB) */
  uint32_t queue_paths;      /* paths per wavefront batch of one scene (0 = 2^25) */
  uint32_t blocks_per_cu;    /* persistent traversal blocks per CU (0 = 8) */
  uint32_t max_jobs_per_batch; /* shot-table entries per batch (0 = 256) */
  uint32_t batches_ahead;    /* batches enqueued per scene before waiting (0 = 2, max 3) */
  uint32_t policy;           /* ASH_SERVE_* batch fill order */
  uint32_t pad;
} ash_server_opts;

/* Batch fill order.  OLDEST_GLOBAL (default): each batch takes all it can
 * from its scene's oldest job and backfills the rest from younger ones, and
 * the scenes are fed in the order of their oldest job with at most
 * 2 x queue_paths paths in flight on the device, so the oldest work shares
 * the device with little else; ROUND_ROBIN: each batch of a scene starts at
 * the next job in turn and every scene keeps its batches in flight (short
 * jobs finish early, long ones share the device with every newcomer);
 * OLDEST_FIRST: oldest-first batches, every scene fed. */
enum { ASH_SERVE_OLDEST_GLOBAL = 0, ASH_SERVE_ROUND_ROBIN = 1, ASH_SERVE_OLDEST_FIRST = 2 };

typedef struct ash_job_desc {
  uint32_t scene;   /* id returned by ash_server_add_scene */
  uint32_t width, height;
  int32_t spp;      /* frames 0 .. spp-1 are accumulated */
  GpuCamera camera;
  GpuPushConstantRaytrace pc; /* as for ash_render */
  int32_t depth_channel;      /* AOV channel for first-hit depth, -1 = none */
  uint32_t flags;             /* ASH_JOB_DENOISE or 0 */
} ash_job_desc;

enum { ASH_JOB_QUEUED = 0, ASH_JOB_RUNNING = 1, ASH_JOB_DONE = 2 };
/* Job flag: after its last frame the job's radiance is denoised (guides:
 * pc.diffuseOutChannel, pc.normalOutChannel, depth_channel) into image 9,
 * which ash_server_read then accepts.  Needs ash_server_set_denoiser. */
enum { ASH_JOB_DENOISE = 1 };

typedef struct ash_job_status {
  int32_t state;       /* ASH_JOB_* or a negative ASH_E_* code */
  uint32_t pad;
  double submit_ms;    /* server clock (ms since ash_server_create) */
  double start_ms;     /* first batch enqueued */
  double done_ms;      /* last batch observed complete (and denoised) */
  uint64_t paths;
  double denoise_ms;   /* device time of the denoise step, 0 if none */
} ash_job_status;

typedef struct ash_server_counters {
  uint64_t jobs_submitted, jobs_done;
  uint64_t batches;
  uint64_t paths;
  uint64_t arena_bytes, arena_used_bytes;
  double busy_ms;      /* wall time with at least one batch in flight */
  double uptime_ms;
} ash_server_counters;

int ash_server_create(int device, const ash_server_opts* opts, ash_server** out);
void ash_server_destroy(ash_server* srv);
const char* ash_server_last_error(ash_server* srv);

/* Upload a scene and build its BVH; it stays resident.  Not concurrent with
 * running jobs of other scenes' uploads (call before submitting). */
int ash_server_add_scene(ash_server* srv, const ash_scene_desc* desc, const ash_build_opts* opts,
                         uint32_t* scene_id);

/* Denoiser used for ASH_JOB_DENOISE jobs (owned by the caller; not used
 * elsewhere while the server runs). */
int ash_server_set_denoiser(ash_server* srv, ash_denoiser* d);

/* Queue a job; returns at once with its id. */
int ash_server_submit(ash_server* srv, const ash_job_desc* job, uint64_t* job_id);

/* Wait up to timeout_ms (-1 = forever) for a job; fills status either way.
 * Returns ASH_OK when done, ASH_E_STATE on timeout. */
int ash_server_wait(ash_server* srv, uint64_t job_id, int32_t timeout_ms, ash_job_status* status);

/* Copy output image `image` (0..8, or 9 = denoised) of a finished job to host RGBA32F. */
int ash_server_read(ash_server* srv, uint64_t job_id, int32_t image, float* host_rgba);

/* Forget a finished job and return its images to the arena. */
int ash_server_release(ash_server* srv, uint64_t job_id);

int ash_server_stats(ash_server* srv, ash_server_counters* stats);

#ifdef __cplusplus
}
#endif

#endif /* ASUNA_SERVER_H */
