// Resident render server (include/asuna_server.h): scenes stay in HBM, jobs
// are admitted by an HBM budget and co-batched per scene into wavefront
// batches; each scene's batches run on its own stream, so batches of
// different scenes overlap on the device.  One scheduler thread owns every
// stream; API calls only touch host state under the mutex.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../../include/asuna_server.h"
#include "ctx_internal.h"

using namespace ashd;

namespace {

using Clock = std::chrono::steady_clock;
constexpr int kRing = 4;  // table slots per scene (batches in flight + 1)

// First-fit sub-allocator over one device block (256-byte granules).
struct Arena {
  char* base = nullptr;
  uint64_t size = 0, used = 0;
  std::map<uint64_t, uint64_t> freeList;  // offset -> bytes
  static uint64_t round(uint64_t b) { return (b + 255) & ~uint64_t(255); }
  bool alloc(uint64_t bytes, uint64_t& off) {
    bytes = round(bytes);
    for (auto it = freeList.begin(); it != freeList.end(); ++it) {
      if (it->second < bytes) continue;
      off = it->first;
      uint64_t rest = it->second - bytes;
      freeList.erase(it);
      if (rest) freeList[off + bytes] = rest;
      used += bytes;
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = round(bytes);
    used -= bytes;
    auto it = freeList.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != freeList.end() && it->first + it->second == next->first) {
      it->second += next->second;
      freeList.erase(next);
    }
    if (it != freeList.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        freeList.erase(it);
      }
    }
  }
};

struct Job {
  uint64_t id = 0;
  ash_job_desc d;
  ShotParams sp;
  uint32_t P = 0;
  uint64_t off = 0, bytes = 0;
  bool admitted = false;
  int32_t framesIssued = 0;  // frames whose every pixel has been issued
  uint32_t nextPix = 0;      // pixel progress inside frame framesIssued
  bool issued = false;
  int inflight = 0;
  int32_t state = ASH_JOB_QUEUED;
  double tSubmit = 0, tStart = -1, tDone = -1;
  double denoiseMs = 0;
  uint32_t planes = ASH_NUM_OUTPUT_IMAGES;  // + 1 denoised plane with ASH_JOB_DENOISE
  hipEvent_t dnStart = nullptr, dnDone = nullptr;  // brackets of the job's denoise on the denoiser's stream
  ~Job() {
    if (dnStart) (void)hipEventDestroy(dnStart);
    if (dnDone) (void)hipEventDestroy(dnDone);
  }
};

struct Batch {
  hipEvent_t done;
  std::vector<Job*> jobs;
  uint64_t paths;
};

struct SceneSlot {
  ash_ctx* ctx = nullptr;
  uint32_t grid = 0;
  std::vector<Job*> active;  // admitted jobs with frames left to issue
  size_t rr = 0;             // round-robin start
  std::deque<Batch> inflight;
  WorkUnit* dUnits[kRing] = {};
  ShotParams* dShots[kRing] = {};
  WorkUnit* hUnits[kRing] = {};    // pinned staging
  ShotParams* hShots[kRing] = {};  // pinned staging
  hipEvent_t events[kRing] = {};
  int next = 0;
};

}  // namespace

struct ash_server {
  int device = 0;
  ash_server_opts o{};
  std::string err;
  Arena arena;
  hipStream_t io = nullptr;  // readback stream
  std::vector<SceneSlot*> scenes;
  std::unordered_map<uint64_t, Job*> jobs;
  std::deque<Job*> pending;
  std::vector<Job*> denoising;  // jobs whose denoise is enqueued and not yet observed complete
  std::mutex m;
  std::condition_variable cvWork, cvDone;
  std::thread th;
  bool stop = false;
  uint64_t nextId = 1;
  Clock::time_point t0;
  ash_server_counters st{};
  double busySince = -1;
  int loopError = 0;
  ash_denoiser* den = nullptr;

  double now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

static int sfail(ash_server* s, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  s->err = buf;
  return code;
}

// Fills one batch of `sc` from its active jobs; false when there is no work.
static int launchBatch(ash_server* s, SceneSlot& sc, bool& launched) {
  launched = false;
  if (sc.active.empty()) return ASH_OK;
  const uint32_t Q = s->o.queue_paths;
  const uint32_t maxJobs = s->o.max_jobs_per_batch;
  const int slot = sc.next;
  WorkUnit* units = sc.hUnits[slot];
  ShotParams* shots = sc.hShots[slot];
  uint32_t nu = 0, n = 0, nAcc = 0;
  int maxDepth = 1;
  std::vector<Job*> batchJobs;
  const size_t N = sc.active.size();
  for (size_t k = 0; k < N && n < Q && nu < maxJobs; ++k) {
    Job* j = sc.active[s->o.policy != ASH_SERVE_ROUND_ROBIN ? k : (sc.rr + k) % N];
    uint32_t room = Q - n;
    WorkUnit u;
    std::memset(&u, 0, sizeof u);
    u.shot = nu;
    u.f0 = j->framesIssued;
    int32_t left = j->d.spp - j->framesIssued;
    if (j->nextPix == 0 && j->P <= room) {
      u.pix0 = 0;
      u.np = j->P;
      u.nf = uint32_t(std::min<int64_t>(left, room / j->P));
      j->framesIssued += int32_t(u.nf);
    } else {
      if (room < 4096 && n > 0) break;  // leave small remainders for the next batch
      u.pix0 = j->nextPix;
      u.np = std::min(j->P - j->nextPix, room);
      u.nf = 1;
      j->nextPix += u.np;
      if (j->nextPix == j->P) {
        j->nextPix = 0;
        j->framesIssued += 1;
      }
    }
    u.base = n;
    u.accBase = nAcc;
    n += u.np * u.nf;
    nAcc += u.np;
    units[nu] = u;
    shots[nu] = j->sp;
    ++nu;
    maxDepth = std::max(maxDepth, j->sp.pc.maxPathDepth);
    if (j->framesIssued >= j->d.spp) j->issued = true;
    j->inflight++;
    if (j->tStart < 0) {
      j->tStart = s->now();
      j->state = ASH_JOB_RUNNING;
    }
    batchJobs.push_back(j);
  }
  if (nu == 0) return ASH_OK;
  sc.rr = (sc.rr + 1) % std::max<size_t>(N, 1);
  sc.active.erase(std::remove_if(sc.active.begin(), sc.active.end(), [](Job* j) { return j->issued; }),
                  sc.active.end());
  hipStream_t st = ctxStream(sc.ctx);
  if (hipMemcpyAsync(sc.dUnits[slot], units, nu * sizeof(WorkUnit), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(sc.dShots[slot], shots, nu * sizeof(ShotParams), hipMemcpyHostToDevice, st) != hipSuccess)
    return sfail(s, ASH_E_HIP, "server: table upload failed");
  BatchDesc b;
  std::memset(&b, 0, sizeof b);
  b.n = n;
  b.nAcc = nAcc;
  b.nUnits = nu;
  b.shots = 0;  // always through the table: one captured bounce loop per table slot
  b.units = sc.dUnits[slot];
  b.shotTab = sc.dShots[slot];
  b.one = shots[0];
  ash_render_opts ro;
  std::memset(&ro, 0, sizeof ro);
  ro.depthOutChannel = -1;
  ash_stats bst;
  std::memset(&bst, 0, sizeof bst);
  int rc = ctxRunBatch(sc.ctx, b, maxDepth, sc.grid, &ro, &bst);
  if (rc) return sfail(s, rc, "server: batch failed: %s", ctxError(sc.ctx));
  if (hipEventRecord(sc.events[slot], st) != hipSuccess) return sfail(s, ASH_E_HIP, "server: event record failed");
  sc.inflight.push_back(Batch{sc.events[slot], std::move(batchJobs), n});
  sc.next = (slot + 1) % kRing;
  s->st.batches++;
  s->st.paths += n;
  launched = true;
  return ASH_OK;
}

static void schedulerLoop(ash_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->m);
  const size_t ahead = s->o.batches_ahead;
  while (!s->stop) {
    bool progress = false;
    bool any = false;
    // retire finished batches
    for (SceneSlot* sc : s->scenes) {
      while (!sc->inflight.empty()) {
        hipError_t q = hipEventQuery(sc->inflight.front().done);
        if (q == hipErrorNotReady) break;
        Batch& bt = sc->inflight.front();
        double t = s->now();
        for (Job* j : bt.jobs) {
          if (--j->inflight == 0 && j->issued) {
            j->state = q == hipSuccess ? ASH_JOB_DONE : ASH_E_HIP;
            if (j->state == ASH_JOB_DONE && (j->d.flags & ASH_JOB_DENOISE)) {
              // the batch event has completed: the job's planes are final.  The
              // denoise is enqueued on the denoiser's stream (it overlaps the
              // scenes' batches); the job is done when its event completes.
              const size_t P = j->P;
              auto plane = [&](int32_t ch) -> const void* {
                return ch >= 0 && ch < ASH_NUM_OUTPUT_IMAGES - 1 ? j->sp.images + size_t(ch + 1) * P : nullptr;
              };
              int rc = ASH_OK;
              if ((!j->dnStart && hipEventCreate(&j->dnStart) != hipSuccess) ||
                  (!j->dnDone && hipEventCreate(&j->dnDone) != hipSuccess))
                rc = ASH_E_HIP;
              if (!rc)
                rc = denoisePlanesAsync(s->den, j->sp.images, plane(j->d.pc.diffuseOutChannel),
                                        plane(j->d.pc.normalOutChannel), plane(j->d.depth_channel), j->d.width,
                                        j->d.height, j->sp.images + 9 * P, j->dnStart, j->dnDone);
              if (!rc) {
                j->state = ASH_JOB_RUNNING;
                s->denoising.push_back(j);
                continue;
              }
              j->state = rc;
            }
            j->tDone = s->now();
            s->st.jobs_done++;
          }
        }
        (void)t;
        sc->inflight.pop_front();
        progress = true;
        s->cvDone.notify_all();
      }
    }
    // retire finished denoises
    for (size_t k = 0; k < s->denoising.size();) {
      Job* j = s->denoising[k];
      hipError_t q = hipEventQuery(j->dnDone);
      if (q == hipErrorNotReady) {
        ++k;
        continue;
      }
      float ms = 0;
      if (q == hipSuccess && hipEventElapsedTime(&ms, j->dnStart, j->dnDone) == hipSuccess) {
        j->denoiseMs = ms;
        j->state = ASH_JOB_DONE;
      } else {
        j->state = ASH_E_HIP;
      }
      j->tDone = s->now();
      s->st.jobs_done++;
      s->denoising[k] = s->denoising.back();
      s->denoising.pop_back();
      progress = true;
      s->cvDone.notify_all();
    }
    // admit queued jobs in arrival order while their images fit
    while (!s->pending.empty()) {
      Job* j = s->pending.front();
      if (!s->arena.alloc(j->bytes, j->off)) break;
      SceneSlot* sc = s->scenes[j->d.scene];
      j->admitted = true;
      j->sp.images = reinterpret_cast<float4*>(s->arena.base + j->off);
      (void)hipMemsetAsync(j->sp.images, 0, j->bytes, ctxStream(sc->ctx));
      sc->active.push_back(j);
      s->pending.pop_front();
      progress = true;
    }
    // keep the scenes' streams fed: all of them, or (OLDEST_GLOBAL) in the
    // order of their oldest job while the device holds fewer than
    // 2 x queue_paths paths
    std::vector<SceneSlot*> order;
    uint64_t flying = 0, cap = ~0ull;
    for (SceneSlot* sc : s->scenes) {
      order.push_back(sc);
      for (const Batch& bt : sc->inflight) flying += bt.paths;
    }
    if (s->o.policy == ASH_SERVE_OLDEST_GLOBAL) {
      cap = 2ull * s->o.queue_paths;
      order.erase(std::remove_if(order.begin(), order.end(), [](SceneSlot* sc) { return sc->active.empty(); }),
                  order.end());
      std::stable_sort(order.begin(), order.end(), [](SceneSlot* a, SceneSlot* b) {
        return a->active.front()->tSubmit < b->active.front()->tSubmit;
      });
    }
    for (SceneSlot* sc : order) {
      while (sc->inflight.size() < ahead && flying < cap) {
        bool launched = false;
        int rc = launchBatch(s, *sc, launched);
        if (rc) {
          s->loopError = rc;
          s->stop = true;
          break;
        }
        if (!launched) break;
        flying += sc->inflight.back().paths;
        progress = true;
      }
    }
    for (SceneSlot* sc : s->scenes) any = any || !sc->inflight.empty();
    any = any || !s->denoising.empty();
    double t = s->now();
    if (any && s->busySince < 0) s->busySince = t;
    if (!any && s->busySince >= 0) {
      s->st.busy_ms += t - s->busySince;
      s->busySince = -1;
    }
    if (progress) continue;
    if (any) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(50));
      lk.lock();
    } else {
      s->cvWork.wait(lk);
    }
  }
  // drain
  for (SceneSlot* sc : s->scenes)
    for (Batch& bt : sc->inflight) (void)hipEventSynchronize(bt.done);
  for (Job* j : s->denoising) (void)hipEventSynchronize(j->dnDone);
  s->cvDone.notify_all();
}

extern "C" {

int ash_server_create(int device, const ash_server_opts* opts, ash_server** out) {
  if (!out) return ASH_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return ASH_E_HIP;
  ash_server* s = new ash_server();
  s->device = device;
  if (opts) s->o = *opts;
  if (!s->o.hbm_budget_bytes) s->o.hbm_budget_bytes = 8ull << 30;
  if (!s->o.queue_paths) s->o.queue_paths = 1u << 25;
  if (!s->o.blocks_per_cu) s->o.blocks_per_cu = 8;
  if (!s->o.max_jobs_per_batch) s->o.max_jobs_per_batch = 256;
  if (!s->o.batches_ahead) s->o.batches_ahead = 2;
  s->o.batches_ahead = std::min<uint32_t>(s->o.batches_ahead, kRing - 1);
  if (s->o.queue_paths < 4096 || s->o.queue_paths > (1u << 26) || s->o.blocks_per_cu > 64 ||
      s->o.max_jobs_per_batch > 65535 || s->o.policy > ASH_SERVE_OLDEST_FIRST) {
    delete s;
    return ASH_E_INVALID;
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->io, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&s->arena.base, s->o.hbm_budget_bytes) != hipSuccess) {
    if (s->io) (void)hipStreamDestroy(s->io);
    delete s;
    return ASH_E_NOMEM;
  }
  s->arena.size = s->o.hbm_budget_bytes;
  s->arena.freeList[0] = s->arena.size & ~uint64_t(255);
  s->t0 = Clock::now();
  s->th = std::thread(schedulerLoop, s);
  *out = s;
  return ASH_OK;
}

void ash_server_destroy(ash_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
  }
  s->cvWork.notify_all();
  if (s->th.joinable()) s->th.join();
  (void)hipSetDevice(s->device);
  for (SceneSlot* sc : s->scenes) {
    for (int k = 0; k < kRing; ++k) {
      if (sc->dUnits[k]) (void)hipFree(sc->dUnits[k]);
      if (sc->dShots[k]) (void)hipFree(sc->dShots[k]);
      if (sc->hUnits[k]) (void)hipHostFree(sc->hUnits[k]);
      if (sc->hShots[k]) (void)hipHostFree(sc->hShots[k]);
      if (sc->events[k]) (void)hipEventDestroy(sc->events[k]);
    }
    ash_destroy(sc->ctx);
    delete sc;
  }
  for (auto& kv : s->jobs) delete kv.second;
  if (s->arena.base) (void)hipFree(s->arena.base);
  if (s->io) (void)hipStreamDestroy(s->io);
  delete s;
}

const char* ash_server_last_error(ash_server* s) { return s ? s->err.c_str() : "null server"; }

int ash_server_add_scene(ash_server* s, const ash_scene_desc* desc, const ash_build_opts* bo, uint32_t* id) {
  if (!s || !desc || !id) return ASH_E_INVALID;
  SceneSlot* sc = new SceneSlot();
  int rc = ash_create(s->device, &sc->ctx);
  if (rc) {
    delete sc;
    return sfail(s, rc, "server: ash_create failed");
  }
  rc = ash_upload_scene(sc->ctx, desc);
  if (!rc) rc = ash_build_accel(sc->ctx, bo, nullptr);
  if (!rc) {
    sc->grid = uint32_t(ctxNumCU(sc->ctx)) * s->o.blocks_per_cu;
    rc = ctxPrepare(sc->ctx, s->o.queue_paths, sc->grid);
  }
  if (rc) {
    sfail(s, rc, "server: scene setup failed: %s", ash_last_error(sc->ctx));
    ash_destroy(sc->ctx);
    delete sc;
    return rc;
  }
  const uint32_t mj = s->o.max_jobs_per_batch;
  for (int k = 0; k < kRing; ++k) {
    if (hipMalloc(&sc->dUnits[k], mj * sizeof(WorkUnit)) != hipSuccess ||
        hipMalloc(&sc->dShots[k], mj * sizeof(ShotParams)) != hipSuccess ||
        hipHostMalloc(&sc->hUnits[k], mj * sizeof(WorkUnit), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&sc->hShots[k], mj * sizeof(ShotParams), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&sc->events[k], hipEventDisableTiming) != hipSuccess) {
      s->scenes.push_back(sc);  // released by ash_server_destroy
      return sfail(s, ASH_E_NOMEM, "server: table allocation failed");
    }
  }
  std::lock_guard<std::mutex> lk(s->m);
  *id = uint32_t(s->scenes.size());
  s->scenes.push_back(sc);
  return ASH_OK;
}

int ash_server_set_denoiser(ash_server* s, ash_denoiser* den) {
  if (!s) return ASH_E_INVALID;
  std::lock_guard<std::mutex> lk(s->m);
  s->den = den;
  return ASH_OK;
}

int ash_server_submit(ash_server* s, const ash_job_desc* d, uint64_t* id) {
  if (!s || !d || !id) return ASH_E_INVALID;
  std::lock_guard<std::mutex> lk(s->m);
  if (s->loopError) return sfail(s, s->loopError, "server stopped: %s", s->err.c_str());
  if (d->scene >= s->scenes.size()) return sfail(s, ASH_E_INVALID, "unknown scene %u", d->scene);
  if (d->width == 0 || d->height == 0 || d->spp <= 0 || uint64_t(d->width) * d->height > (1u << 28))
    return sfail(s, ASH_E_INVALID, "bad job size");
  if (d->depth_channel >= int32_t(ASH_NUM_OUTPUT_IMAGES - 1)) return sfail(s, ASH_E_INVALID, "depth channel out of range");
  if (d->flags & ~uint32_t(ASH_JOB_DENOISE)) return sfail(s, ASH_E_INVALID, "unknown job flags");
  if ((d->flags & ASH_JOB_DENOISE) && !s->den) return sfail(s, ASH_E_STATE, "denoise flag without a denoiser");
  int rc = ctxCheckPc(s->scenes[d->scene]->ctx, &d->pc);
  if (rc) return sfail(s, rc, "bad push constants: %s", ctxError(s->scenes[d->scene]->ctx));
  Job* j = new Job();
  j->id = s->nextId++;
  j->d = *d;
  j->P = d->width * d->height;
  j->planes = ASH_NUM_OUTPUT_IMAGES + ((d->flags & ASH_JOB_DENOISE) ? 1 : 0);
  j->bytes = uint64_t(j->P) * j->planes * sizeof(float4);
  if (Arena::round(j->bytes) > s->arena.size) {
    delete j;
    return sfail(s, ASH_E_NOMEM, "job images exceed the HBM budget");
  }
  j->sp = ctxMakeShot(&d->camera, &d->pc, d->depth_channel, nullptr, d->width, d->height);
  j->tSubmit = s->now();
  s->jobs[j->id] = j;
  s->pending.push_back(j);
  s->st.jobs_submitted++;
  *id = j->id;
  s->cvWork.notify_one();
  return ASH_OK;
}

static void fillStatus(const Job* j, ash_job_status* st) {
  if (!st) return;
  std::memset(st, 0, sizeof *st);
  st->state = j->state;
  st->submit_ms = j->tSubmit;
  st->start_ms = j->tStart;
  st->done_ms = j->tDone;
  st->paths = uint64_t(j->P) * uint64_t(j->d.spp);
  st->denoise_ms = j->denoiseMs;
}

int ash_server_wait(ash_server* s, uint64_t id, int32_t timeout_ms, ash_job_status* st) {
  if (!s) return ASH_E_INVALID;
  std::unique_lock<std::mutex> lk(s->m);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return sfail(s, ASH_E_INVALID, "unknown job %llu", (unsigned long long)id);
  Job* j = it->second;
  auto doneOrStopped = [&] { return j->state == ASH_JOB_DONE || j->state < 0 || s->stop; };
  if (timeout_ms < 0) s->cvDone.wait(lk, doneOrStopped);
  else s->cvDone.wait_for(lk, std::chrono::milliseconds(timeout_ms), doneOrStopped);
  fillStatus(j, st);
  if (j->state < 0) return j->state;
  if (j->state != ASH_JOB_DONE) return s->loopError ? s->loopError : ASH_E_STATE;
  return ASH_OK;
}

int ash_server_read(ash_server* s, uint64_t id, int32_t image, float* out) {
  if (!s || !out || image < 0 || image > ASH_NUM_OUTPUT_IMAGES) return ASH_E_INVALID;
  const float4* src;
  size_t plane;
  {
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return sfail(s, ASH_E_INVALID, "unknown job");
    Job* j = it->second;
    if (j->state != ASH_JOB_DONE) return sfail(s, ASH_E_STATE, "job not finished");
    if (uint32_t(image) >= j->planes) return sfail(s, ASH_E_INVALID, "image %d not kept for this job", image);
    plane = j->P;
    src = j->sp.images + size_t(image) * plane;
  }
  if (hipMemcpyAsync(out, src, plane * sizeof(float4), hipMemcpyDeviceToHost, s->io) != hipSuccess ||
      hipStreamSynchronize(s->io) != hipSuccess)
    return sfail(s, ASH_E_HIP, "readback failed");
  return ASH_OK;
}

int ash_server_release(ash_server* s, uint64_t id) {
  if (!s) return ASH_E_INVALID;
  std::lock_guard<std::mutex> lk(s->m);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return sfail(s, ASH_E_INVALID, "unknown job");
  Job* j = it->second;
  if (j->state != ASH_JOB_DONE && j->state >= 0) return sfail(s, ASH_E_STATE, "job still running");
  if (j->admitted) s->arena.release(j->off, j->bytes);
  s->jobs.erase(it);
  delete j;
  s->cvWork.notify_one();
  return ASH_OK;
}

int ash_server_stats(ash_server* s, ash_server_counters* out) {
  if (!s || !out) return ASH_E_INVALID;
  std::lock_guard<std::mutex> lk(s->m);
  *out = s->st;
  out->arena_bytes = s->arena.size;
  out->arena_used_bytes = s->arena.used;
  out->uptime_ms = s->now();
  if (s->busySince >= 0) out->busy_ms += s->now() - s->busySince;
  return ASH_OK;
}

}  // extern "C"
