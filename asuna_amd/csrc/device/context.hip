// C ABI of the render library (include/asuna_hip.h): device context, scene
// upload, BVH build, wavefront render loop, output readback.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../host/scene.h"
#include "ctx_internal.h"

using namespace ashd;

static_assert(sizeof(GpuVertex) == 44, "GpuVertex must be 44 bytes (vertex.h)");
static_assert(sizeof(GpuMaterial) == 132, "GpuMaterial must be 132 bytes (material.h)");
static_assert(sizeof(GpuLight) == 76, "GpuLight must be 76 bytes (light.h)");
static_assert(sizeof(GpuCamera) == 224, "GpuCamera must be 224 bytes (camera.h)");
static_assert(sizeof(GpuPushConstantRaytrace) == 84, "push constants must be 84 bytes (pushconstant.h)");
static_assert(sizeof(GpuSunAndSky) == 96, "GpuSunAndSky must be 96 bytes (sun_and_sky.h)");
static_assert(offsetof(GpuSunAndSky, in_use) == 92, "GpuSunAndSky.in_use offset");
static_assert(offsetof(GpuMaterial, type) == 128, "GpuMaterial.type offset");
static_assert(offsetof(GpuLight, doubleSide) == 72, "GpuLight.doubleSide offset");
static_assert(offsetof(GpuCamera, type) == 208, "GpuCamera.type offset");
static_assert(offsetof(GpuPushConstantRaytrace, uvOutChannel) == 80, "push constant uvOutChannel offset");

namespace {

template <class T> struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    return hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
  }
};

}  // namespace

// Identity of a captured bounce loop: every kernel argument it recorded
// (the shading kernel reads only the shot table / single-shot copy of the
// batch descriptor).
struct GraphKey {
  SceneView sc;
  Wavefront w;
  const ShotParams* tab;
  ShotParams one;
  uint32_t grid, iters, n, single, lds, scheme;
};

struct GraphEntry {
  GraphKey key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t lastUse = 0;
};

static void releaseGraph(GraphEntry& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g.exec = nullptr;
  g.graph = nullptr;
}

struct ash_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int numCU = 0;
  // scene
  bool hasScene = false, hasAccel = false;
  DBuf<DevInstance> inst;
  DBuf<float4> verts;
  DBuf<uint32_t> idx, instTriOffset;
  DBuf<DevMaterial> mats;
  DBuf<GpuLight> lights;
  std::vector<DBuf<float4>> texData;
  DBuf<DevImage> texTab;
  DBuf<float4> env;
  DBuf<float2> marg;
  DBuf<float> condX;  // inverse-CDF channel of the conditional table (its pdf channel is the env texels' alpha)
  uint32_t envW = 1, envH = 1;
  bool hasEnv = false;
  DBuf<DevSky> sky;
  uint32_t numTris = 0, numInst = 0;
  uint32_t maxDepth = 0;
  float4* nodes = nullptr;
  float4* nodes4 = nullptr;  // 4-wide tree when built with branching 4
  float4* nodesQ = nullptr;  // its quantized 64-byte copy (node format 2)
  float4* tris = nullptr;
  float4* shadeTris = nullptr;  // per leaf-ordered triangle: its three object-space vertex pairs
  uint32_t stackNeed = 0;    // traversal stack entries the current tree can need
  uint64_t nodeCount = 0;
  // film (single shot)
  uint32_t width = 0, height = 0;
  DBuf<float4> images;
  DBuf<ShotParams> shotTab;
  // multi-shot outputs
  DBuf<float4> shotImages;
  uint32_t shotCount = 0, shotW = 0, shotH = 0;
  // wavefront
  Wavefront wf{};
  DBuf<float4> wRayO0, wRayO1, wRayD0, wRayD1, wThr0, wThr1, wRad, wLd, wHitF, wShO, wShD;
  DBuf<uint4> wInfo0, wInfo1;
  DBuf<int4> wHitI;
  DBuf<int> wSide, wSpill;
  DBuf<float> wSpillT;
  DBuf<uint32_t> wCtr;
  DBuf<unsigned long long> wCnt64;
  DBuf<WorkUnit> unitBuf;
  GraphEntry graphs[4];
  uint64_t graphClock = 0;
  bool graphsBroken = false;
  // two-level acceleration: BLAS arrays owned here, TLAS and tables
  std::vector<DevInstance> hostInst;  // host copy of the instance records
  bool twoLevel = false;
  std::vector<float4*> blasArrays;
  DBuf<float4> tlas;
  DBuf<uint32_t> tlasInst, instBlas;
  DBuf<DevBlas> blasTab;
  // device address of this GPU's allocation-free render scratch (null when the lookup failed)
  unsigned char* scratch = nullptr;
  hipError_t scratchErr = hipSuccess;

  void releaseTwoLevel() {
    for (float4* p : blasArrays) (void)hipFree(p);
    blasArrays.clear();
    tlas.release(), tlasInst.release(), instBlas.release(), blasTab.release();
    twoLevel = false;
  }
  ~ash_ctx() {
    releaseTwoLevel();
    if (nodes) (void)hipFree(nodes);
    if (nodes4) (void)hipFree(nodes4);
    if (nodesQ) (void)hipFree(nodesQ);
    if (tris) (void)hipFree(tris);
    if (shadeTris) (void)hipFree(shadeTris);
    for (auto& t : texData) t.release();
    for (auto* b : {&verts, &env, &images, &shotImages, &wRayO0, &wRayO1, &wRayD0, &wRayD1, &wThr0, &wThr1, &wRad, &wLd,
                    &wHitF, &wShO, &wShD})
      b->release();
    inst.release(), idx.release(), instTriOffset.release(), mats.release(), lights.release(), texTab.release();
    marg.release(), condX.release(), shotTab.release(), wInfo0.release(), wInfo1.release(), wHitI.release();
    wSide.release(), wSpill.release(), wSpillT.release(), wCtr.release(), wCnt64.release();
    unitBuf.release(), sky.release();
    for (auto& g : graphs) releaseGraph(g);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static int fail(ash_ctx* c, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPC(ctx, x)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(ctx, ASH_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

#define MALLOC(ctx, buf, n)                                                                                 \
  do {                                                                                                      \
    hipError_t e_ = (buf).alloc(n);                                                                         \
    if (e_ != hipSuccess) return fail(ctx, ASH_E_NOMEM, "device allocation of %zu elements failed: %s", size_t(n), \
                                      hipGetErrorString(e_));                                               \
  } while (0)

// Affine inverse in double, rounded once (gl_WorldToObjectEXT).
static void invert3x4(const float* m, float* out) {
  double a[3][3], t[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) a[r][c] = m[r * 4 + c];
    t[r] = m[r * 4 + 3];
  }
  double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
               a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  double inv[3][3];
  inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
  inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
  inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
  inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
  inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
  inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
  inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
  inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
  inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  for (int r = 0; r < 3; ++r) {
    double tt = -(inv[r][0] * t[0] + inv[r][1] * t[1] + inv[r][2] * t[2]);
    for (int c = 0; c < 3; ++c) out[r * 4 + c] = float(inv[r][c]);
    out[r * 4 + 3] = float(tt);
  }
}

static SceneView viewOf(const ash_ctx* c) {
  SceneView v;
  std::memset(&v, 0, sizeof v);
  v.nodes = c->nodes;
  v.nodes4 = c->nodes4;
  v.nodesQ = c->nodesQ;
  v.tris = c->tris;
  v.shadeTris = c->shadeTris;
  v.inst = c->inst.p;
  v.verts = c->verts.p;
  v.idx = c->idx.p;
  v.mats = c->mats.p;
  v.numMats = uint32_t(c->mats.n);
  v.lights = c->lights.p;
  v.textures = c->texTab.p;
  v.numTextures = uint32_t(c->texTab.n);
  v.hasEnv = c->hasEnv ? 1u : 0u;
  v.env.px = c->env.p;
  v.env.w = c->envW;
  v.env.h = c->envH;
  v.marg = c->marg.p;
  v.condX = c->condX.p;
  v.sky = c->sky.p;
  v.twoLevel = c->twoLevel ? 1u : 0u;
  v.tlas = c->tlas.p;
  v.tlasInst = c->tlasInst.p;
  v.blas = c->blasTab.p;
  v.instBlas = c->instBlas.p;
  return v;
}

static int ensureWavefront(ash_ctx* c, uint32_t cap) {
  if (c->wf.capacity >= cap && c->wRayO0.p) return ASH_OK;
  MALLOC(c, c->wRayO0, cap);
  MALLOC(c, c->wRayO1, cap);
  MALLOC(c, c->wRayD0, cap);
  MALLOC(c, c->wRayD1, cap);
  MALLOC(c, c->wThr0, cap);
  MALLOC(c, c->wThr1, cap);
  MALLOC(c, c->wRad, cap);
  MALLOC(c, c->wLd, cap);
  MALLOC(c, c->wHitF, cap);
  MALLOC(c, c->wShO, cap);
  MALLOC(c, c->wShD, cap);
  MALLOC(c, c->wInfo0, cap);
  MALLOC(c, c->wInfo1, cap);
  MALLOC(c, c->wHitI, cap);
  MALLOC(c, c->wSide, cap);
  MALLOC(c, c->wCtr, kNumCtr);
  MALLOC(c, c->wCnt64, kNumCnt64);
  Wavefront& w = c->wf;
  w.capacity = cap;
  w.rayO[0] = c->wRayO0.p;
  w.rayO[1] = c->wRayO1.p;
  w.rayD[0] = c->wRayD0.p;
  w.rayD[1] = c->wRayD1.p;
  w.thr[0] = c->wThr0.p;
  w.thr[1] = c->wThr1.p;
  w.rad = c->wRad.p;
  w.ld = c->wLd.p;
  w.info[0] = c->wInfo0.p;
  w.info[1] = c->wInfo1.p;
  w.hitF = c->wHitF.p;
  w.hitI = c->wHitI.p;
  w.side = c->wSide.p;
  w.shO = c->wShO.p;
  w.shD = c->wShD.p;
  w.ctr = c->wCtr.p;
  w.cnt64 = c->wCnt64.p;
  HIPC(c, hipMemsetAsync(c->wCtr.p, 0, kNumCtr * sizeof(uint32_t), c->stream));
  HIPC(c, hipMemsetAsync(c->wCnt64.p, 0, kNumCnt64 * sizeof(unsigned long long), c->stream));
  return ASH_OK;
}

// Global spill space for the traversal stacks of all persistent threads,
// sized to the deepest stack the current tree can produce.
static int ensureSpill(ash_ctx* c, uint32_t grid) {
  uint32_t cap =
      std::max<uint32_t>(uint32_t(kSpill), c->stackNeed > uint32_t(kStackLdsMin) ? c->stackNeed - kStackLdsMin : 0);
  uint32_t threads = grid * kTraceBlock;
  Wavefront& w = c->wf;
  if (w.spill && w.spillCap >= cap && w.spillThreads == threads) return ASH_OK;
  MALLOC(c, c->wSpill, size_t(threads) * cap);
  MALLOC(c, c->wSpillT, size_t(threads) * cap);
  w.spill = c->wSpill.p;
  w.spillT = c->wSpillT.p;
  w.spillThreads = threads;
  w.spillCap = cap;
  return ASH_OK;
}

static int checkPc(ash_ctx* c, const GpuPushConstantRaytrace* pc) {
  if (pc->numLights < 0 || size_t(pc->numLights) + 1 > c->lights.n)
    return fail(c, ASH_E_INVALID, "numLights = %d but %zu lights (incl. dummy) were uploaded", pc->numLights, c->lights.n);
  if (pc->nMultiChannel > ASH_NUM_OUTPUT_IMAGES - 1) return fail(c, ASH_E_INVALID, "nMultiChannel > 8");
  const int32_t chans[] = {pc->diffuseOutChannel,  pc->specularOutChannel, pc->roughnessOutChannel, pc->normalOutChannel,
                           pc->positionOutChannel, pc->tangentOutChannel,  pc->uvOutChannel};
  for (int32_t ch : chans)
    if (ch >= int32_t(ASH_NUM_OUTPUT_IMAGES - 1)) return fail(c, ASH_E_INVALID, "AOV channel index %d out of range", ch);
  if (pc->hasEnvMap == 1 && !c->hasEnv) return fail(c, ASH_E_INVALID, "hasEnvMap = 1 but no env map was uploaded");
  if (pc->maxPathDepth > 0xfffe) return fail(c, ASH_E_INVALID, "maxPathDepth too large");
  return ASH_OK;
}

// The bounce loop of a batch.  On the default traversal scheme the shadow
// rays of each bounce are traced in the same persistent launch as the
// extension rays of the next one (extend, then per bounce: shade, fused
// trace; the last bounce's shadow rays alone); otherwise each bounce is
// extend, shade, shadow.  `shadeRays` sizes the shading grid.
// `count` selects the counting kernels; with `st` set, every traversal and
// shading launch is bracketed by HIP events and its time added to st.
static int launchLoop(ash_ctx* c, const Wavefront& wf, const SceneView& sc, const BatchDesc& b, uint32_t iters,
                      uint32_t grid, uint32_t lds, uint32_t scheme, uint32_t shadeRays, bool count, ash_stats* st) {
  hipStream_t s = c->stream;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (st) {
    HIPC(c, hipEventCreate(&ev[0]));
    HIPC(c, hipEventCreate(&ev[1]));
  }
  int rc = ASH_OK;
  auto timed = [&](double* acc, auto&& fn) {
    if (rc != ASH_OK) return;
    if (!st) {
      fn();
      return;
    }
    float ms = 0;
    if (hipEventRecord(ev[0], s) != hipSuccess) rc = ASH_E_HIP;
    fn();
    if (rc == ASH_OK && (hipEventRecord(ev[1], s) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                         hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
      rc = ASH_E_HIP;
    *acc += ms;
  };
  double scratch = 0;
  double* tExt = st ? &st->extend_ms : &scratch;
  double* tShade = st ? &st->shade_ms : &scratch;
  double* tShadow = st ? &st->shadow_ms : &scratch;
  double* tTrace = st ? &st->trace_ms : &scratch;
  int cur = 0;
  const bool coarse = c->numTris <= kCoarseFetchTris;
  if (!canFuseTrace(sc, scheme)) {
    for (uint32_t it = 0; it < iters; ++it) {
      launchResetCounters(wf, cur, s);
      timed(tExt, [&] {
        if (it == 0) launchExtendCamera(sc, wf, lds, scheme, grid, count, coarse, s);
        else launchExtend(sc, wf, cur, lds, scheme, grid, count, s);
      });
      timed(tShade, [&] { launchShade(sc, b, wf, cur, shadeRays, s); });
      timed(tShadow, [&] { launchShadow(sc, wf, scheme, grid, count, s); });
      cur = 1 - cur;
    }
  } else {
    launchResetCounters(wf, 0, s);
    timed(tExt, [&] { launchExtendCamera(sc, wf, lds, scheme, grid, count, coarse, s); });
    for (uint32_t it = 0; it < iters; ++it) {
      if (it > 0) launchResetShade(wf, cur, s);
      timed(tShade, [&] { launchShade(sc, b, wf, cur, shadeRays, s); });
      if (it + 1 < iters) {
        launchResetExtend(wf, s);
        timed(tTrace, [&] { launchTrace(sc, wf, 1 - cur, lds, grid, count, s); });
      } else {
        timed(tShadow, [&] { launchShadow(sc, wf, scheme, grid, count, s); });
      }
      cur = 1 - cur;
    }
  }
  if (st) {
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  return rc;
}

// The shading grid covers the wavefront capacity (threads past the live
// count exit), so one captured loop serves every batch size.
static void launchBounces(ash_ctx* c, const SceneView& sc, const BatchDesc& b, uint32_t iters, uint32_t grid,
                          uint32_t lds, uint32_t scheme) {
  (void)launchLoop(c, c->wf, sc, b, iters, grid, lds, scheme, c->wf.capacity, false, nullptr);
}

// Replays the bounce loop of a batch from a captured HIP graph (captured on
// first use of a key; a few keys are cached).  The loop holds kernel nodes
// only.  Returns false when capture is unavailable (eager launch instead).
static bool graphBounces(ash_ctx* c, const SceneView& sc, const BatchDesc& b, uint32_t iters, uint32_t grid,
                         uint32_t lds, uint32_t scheme) {
  if (c->graphsBroken) return false;
  GraphKey k;
  std::memset(&k, 0, sizeof k);
  std::memcpy(&k.sc, &sc, sizeof sc);
  std::memcpy(&k.w, &c->wf, sizeof c->wf);
  k.tab = b.shotTab;
  if (b.shots == 1) std::memcpy(&k.one, &b.one, sizeof b.one);
  k.grid = grid;
  k.iters = iters;
  k.n = c->wf.capacity;
  k.single = b.shots == 1;
  k.lds = lds;
  k.scheme = scheme;
  ++c->graphClock;
  for (auto& g : c->graphs)
    if (g.exec && std::memcmp(&g.key, &k, sizeof k) == 0) {
      g.lastUse = c->graphClock;
      return hipGraphLaunch(g.exec, c->stream) == hipSuccess;
    }
  GraphEntry* slot = &c->graphs[0];
  for (auto& g : c->graphs)
    if (!g.exec || g.lastUse < slot->lastUse) slot = &g;
  // the evicted graph may still be executing on the stream (launched for an
  // earlier batch): it is destroyed only once the stream has drained
  if (slot->exec && hipStreamSynchronize(c->stream) != hipSuccess) {
    c->graphsBroken = true;
    return false;
  }
  releaseGraph(*slot);
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    c->graphsBroken = true;
    return false;
  }
  launchBounces(c, sc, b, iters, grid, lds, scheme);
  hipGraph_t graph = nullptr;
  if (hipStreamEndCapture(c->stream, &graph) != hipSuccess || !graph) {
    c->graphsBroken = true;
    (void)hipGetLastError();
    return false;
  }
  hipGraphExec_t exec = nullptr;
  if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    c->graphsBroken = true;
    return false;
  }
  slot->key = k;
  slot->graph = graph;
  slot->exec = exec;
  slot->lastUse = c->graphClock;
  return hipGraphLaunch(exec, c->stream) == hipSuccess;
}

// One batch: ray generation, the bounce loop, the ordered accumulation.
static int runBatch(ash_ctx* c, const BatchDesc& b, int maxDepth, uint32_t grid, const ash_render_opts* o,
                    ash_stats* st) {
  SceneView sc = viewOf(c);
  hipStream_t s = c->stream;
  bool timing = o && o->timing;
  bool count = o && o->count_traversal;
  bool eager = timing || count || (o && o->no_graphs);
  uint32_t iters = o && o->max_bounce_iters ? o->max_bounce_iters : uint32_t(std::max(maxDepth, 1));
  uint32_t lds = o && o->lds_stack == uint32_t(kStackLdsMin) ? uint32_t(kStackLdsMin) : uint32_t(kStackLds);
  uint32_t scheme = o ? o->traversal : uint32_t(kSchemeXcdSlices);
  launchRaygen(b, c->wf, s);
  st->batches++;
  if (canFuseTrace(sc, scheme)) {
    st->extend_launches += 1;
    st->shadow_launches += 1;
    st->trace_launches += iters - 1;
  } else {
    st->extend_launches += iters;
    st->shadow_launches += iters;
  }
  if (!eager && graphBounces(c, sc, b, iters, grid, lds, scheme)) {
    st->graph_launches++;
  } else {
    int rc = launchLoop(c, c->wf, sc, b, iters, grid, lds, scheme, b.n, count, timing ? st : nullptr);
    if (rc) return fail(c, rc, "bounce loop: HIP event failure");
  }
  launchAccumulate(b, c->wf, s);
  HIPC(c, hipGetLastError());
  return ASH_OK;
}

// Splits shots x frames into wavefront batches of at most Q paths.  The
// units of a batch cover disjoint (shot, pixel) sets and frames are the
// outer loop, so every pixel folds its frames in increasing order.
static void planBatches(uint32_t nShots, uint32_t P, int32_t first, int32_t nFrames, uint64_t Q,
                        std::vector<std::vector<WorkUnit>>& out) {
  out.clear();
  auto unit = [](uint32_t s, uint32_t pix0, uint32_t np, int32_t f0, uint32_t nf) {
    WorkUnit u;
    std::memset(&u, 0, sizeof u);
    u.shot = s, u.pix0 = pix0, u.np = np, u.f0 = f0, u.nf = nf;
    return u;
  };
  Q = std::max<uint64_t>(Q, 64);
  const int32_t end = first + nFrames;
  uint64_t perFrame = uint64_t(P) * nShots;
  if (perFrame <= Q) {
    int32_t F = int32_t(std::min<uint64_t>(uint64_t(nFrames), Q / perFrame));
    for (int32_t f0 = first; f0 < end; f0 += F) {
      uint32_t nf = uint32_t(std::min(F, end - f0));
      out.emplace_back();
      for (uint32_t s = 0; s < nShots; ++s) out.back().push_back(unit(s, 0, P, f0, nf));
    }
  } else if (P <= Q) {
    uint32_t g = uint32_t(Q / P);
    for (int32_t f = first; f < end; ++f)
      for (uint32_t s0 = 0; s0 < nShots; s0 += g) {
        out.emplace_back();
        for (uint32_t s = s0; s < std::min(nShots, s0 + g); ++s) out.back().push_back(unit(s, 0, P, f, 1));
      }
  } else {
    uint32_t run = uint32_t(Q & ~uint64_t(63));
    for (int32_t f = first; f < end; ++f)
      for (uint32_t s = 0; s < nShots; ++s)
        for (uint32_t p0 = 0; p0 < P; p0 += run) out.push_back({unit(s, p0, std::min(run, P - p0), f, 1)});
  }
  for (auto& bt : out) {
    uint32_t base = 0, acc = 0;
    for (auto& u : bt) {
      u.base = base;
      u.accBase = acc;
      base += u.np * u.nf;
      acc += u.np;
    }
  }
}

static int renderCommon(ash_ctx* c, uint32_t nShots, const ShotParams* hostShots, int32_t firstFrame, int32_t nFrames,
                        const ash_render_opts* o, ash_stats* st) {
  if (!c->hasScene || !c->hasAccel) return fail(c, ASH_E_STATE, "render called before upload_scene/build_accel");
  if (nFrames <= 0 || firstFrame < 0) return fail(c, ASH_E_INVALID, "bad frame range");
  if (nShots == 0 || nShots > 0xffff) return fail(c, ASH_E_INVALID, "1..65535 shots per render");
  uint32_t P = hostShots[0].width * hostShots[0].height;
  int maxDepth = 0;
  for (uint32_t s = 0; s < nShots; ++s) {
    int rc = checkPc(c, &hostShots[s].pc);
    if (rc) return rc;
    maxDepth = std::max(maxDepth, hostShots[s].pc.maxPathDepth);
    if (hostShots[s].width * hostShots[s].height != P) return fail(c, ASH_E_INVALID, "shots must share one film size");
  }
  uint64_t Q = o && o->queue_paths ? o->queue_paths : (1u << 25);
  if (Q > (1u << 30)) return fail(c, ASH_E_INVALID, "queue_paths above 2^30");
  uint32_t bpc = o && o->blocks_per_cu ? o->blocks_per_cu : 8;
  if (bpc > 64) return fail(c, ASH_E_INVALID, "blocks_per_cu above 64");
  if (o && o->lds_stack && o->lds_stack != uint32_t(kStackLdsMin) && o->lds_stack != uint32_t(kStackLds))
    return fail(c, ASH_E_INVALID, "lds_stack must be 0, %d or %d", kStackLdsMin, kStackLds);
  if (o && o->traversal > uint32_t(kSchemeRefill)) return fail(c, ASH_E_INVALID, "traversal must be 0, 1 or 2");
  std::vector<std::vector<WorkUnit>> plan;
  planBatches(nShots, P, firstFrame, nFrames, Q, plan);
  uint32_t cap = 0;
  size_t nUnits = 0;
  for (auto& bt : plan) {
    cap = std::max(cap, bt.back().base + bt.back().np * bt.back().nf);
    nUnits += bt.size();
  }
  int rc = ensureWavefront(c, cap);
  if (rc) return rc;
  uint32_t grid = uint32_t(std::max(1, c->numCU)) * bpc;
  rc = ensureSpill(c, grid);
  if (rc) return rc;
  std::vector<WorkUnit> flat;
  flat.reserve(nUnits);
  for (auto& bt : plan) flat.insert(flat.end(), bt.begin(), bt.end());
  if (c->unitBuf.n < flat.size()) MALLOC(c, c->unitBuf, flat.size());
  if (c->shotTab.n < nShots) MALLOC(c, c->shotTab, nShots);
  HIPC(c, hipStreamSynchronize(c->stream));  // previous users of the tables are done
  HIPC(c, hipMemcpy(c->unitBuf.p, flat.data(), flat.size() * sizeof(WorkUnit), hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->shotTab.p, hostShots, nShots * sizeof(ShotParams), hipMemcpyHostToDevice));
  ash_stats local{};
  ash_stats* S = st ? st : &local;
  std::memset(S, 0, sizeof *S);
  HIPC(c, hipMemsetAsync(c->wf.cnt64, 0, kNumCnt64 * sizeof(unsigned long long), c->stream));
  hipEvent_t e0, e1;
  HIPC(c, hipEventCreate(&e0));
  HIPC(c, hipEventCreate(&e1));
  HIPC(c, hipEventRecord(e0, c->stream));
  size_t off = 0;
  for (auto& bt : plan) {
    BatchDesc b;
    std::memset(&b, 0, sizeof b);
    b.n = bt.back().base + bt.back().np * bt.back().nf;
    b.nAcc = bt.back().accBase + bt.back().np;
    b.nUnits = uint32_t(bt.size());
    b.units = c->unitBuf.p + off;
    b.shotTab = c->shotTab.p;
    // kernel-argument shot constants only for single-shot renders: with
    // several shots every batch reads the table, so one captured bounce
    // loop serves all batches of the same size
    b.shots = nShots;
    b.one = hostShots[0];
    off += bt.size();
    rc = runBatch(c, b, maxDepth, grid, o, S);
    if (rc) break;
  }
  HIPC(c, hipEventRecord(e1, c->stream));
  HIPC(c, hipEventSynchronize(e1));
  float ms = 0;
  HIPC(c, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  unsigned long long cnt[kNumCnt64];
  HIPC(c, hipMemcpy(cnt, c->wf.cnt64, sizeof cnt, hipMemcpyDeviceToHost));
  S->render_ms = ms;
  S->paths = uint64_t(P) * nShots * uint64_t(nFrames);
  S->extension_rays = cnt[4];
  S->shadow_rays = cnt[5];
  S->dielectric_shadow_rays = cnt[6];
  if (o && o->count_traversal) {
    S->ext_node_visits = cnt[0];
    S->ext_tri_tests = cnt[1];
    S->shadow_node_visits = cnt[2];
    S->shadow_tri_tests = cnt[3];
    S->trace_ext_node_visits = cnt[7];
    S->trace_ext_tri_tests = cnt[8];
    S->trace_shadow_node_visits = cnt[9];
    S->trace_shadow_tri_tests = cnt[10];
    S->trace_ext_rays = cnt[11];
    S->trace_shadow_rays = cnt[12];
  }
  return ASH_OK;
}

// ------------------------------------------------- two-level acceleration
// TLAS over the instances' world boxes: a binary tree built on the host by
// median splits of the box centres along their widest axis (the instance
// count is small beside the triangle count), in the BVH2 node layout the
// binary walk reads: child boxes in the parent, leaf = one instance.
struct HostBox {
  double lo[3], hi[3];
};
static float refBits(int32_t r) {
  float f;
  std::memcpy(&f, &r, 4);
  return f;
}
static int32_t bitsRef(float f) {
  int32_t r;
  std::memcpy(&r, &f, 4);
  return r;
}
static void tlasNode(std::vector<float4>& nodes, size_t at, const HostBox& L, int32_t rl, const HostBox& R, int32_t rr) {
  auto f = [](double v, bool up) {
    float x = float(v);
    if (up ? double(x) < v : double(x) > v) x = std::nextafter(x, up ? INFINITY : -INFINITY);
    return x;
  };
  nodes[at + 0] = make_float4(f(L.lo[0], false), f(L.hi[0], true), f(L.lo[1], false), f(L.hi[1], true));
  nodes[at + 1] = make_float4(f(R.lo[0], false), f(R.hi[0], true), f(R.lo[1], false), f(R.hi[1], true));
  nodes[at + 2] = make_float4(f(L.lo[2], false), f(L.hi[2], true), f(R.lo[2], false), f(R.hi[2], true));
  nodes[at + 3] = make_float4(refBits(rl), refBits(rr), 0.f, 0.f);
}
static int32_t tlasBuild(std::vector<float4>& nodes, std::vector<uint32_t>& order, const std::vector<HostBox>& box,
                         std::vector<uint32_t>& items, size_t b, size_t e, HostBox& out, uint32_t depth,
                         uint32_t& maxDepth) {
  maxDepth = std::max(maxDepth, depth);
  if (e - b == 1) {
    out = box[items[b]];
    order.push_back(items[b]);
    return encodeLeaf(uint32_t(order.size() - 1), 1);
  }
  const size_t at = nodes.size();
  nodes.resize(at + 4);
  double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t k = b; k < e; ++k)
    for (int a = 0; a < 3; ++a) {
      const double cc = 0.5 * (box[items[k]].lo[a] + box[items[k]].hi[a]);
      clo[a] = std::min(clo[a], cc), chi[a] = std::max(chi[a], cc);
    }
  int ax = 0;
  for (int a = 1; a < 3; ++a)
    if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
  const size_t m = b + (e - b) / 2;
  std::nth_element(items.begin() + b, items.begin() + m, items.begin() + e, [&](uint32_t x, uint32_t y) {
    const double cx = box[x].lo[ax] + box[x].hi[ax], cy = box[y].lo[ax] + box[y].hi[ax];
    return cx < cy || (cx == cy && x < y);
  });
  HostBox L, R;
  const int32_t rl = tlasBuild(nodes, order, box, items, b, m, L, depth + 1, maxDepth);
  const int32_t rr = tlasBuild(nodes, order, box, items, m, e, R, depth + 1, maxDepth);
  tlasNode(nodes, at, L, rl, R, rr);
  for (int a = 0; a < 3; ++a) out.lo[a] = std::min(L.lo[a], R.lo[a]), out.hi[a] = std::max(L.hi[a], R.hi[a]);
  return int32_t(at / 4);
}

// One object-space BLAS per distinct mesh (instances share a mesh when
// they share its vertex and index ranges), each built by buildBvh over a
// single identity instance, then the TLAS.
static int buildTwoLevel(ash_ctx* c, const BuildArgs& tmpl, ash_build_info* info) {
  hipStream_t s = c->stream;
  std::vector<uint32_t> instBlas(c->numInst, 0xffffffffu);
  std::vector<DevInstance> meshInst;  // the identity instance of each distinct mesh
  for (uint32_t i = 0; i < c->numInst; ++i) {
    const DevInstance& in = c->hostInst[i];
    if (!in.ntri) continue;
    uint32_t k = 0;
    while (k < meshInst.size() && !(meshInst[k].vbase == in.vbase && meshInst[k].ibase == in.ibase &&
                                     meshInst[k].ntri == in.ntri))
      ++k;
    if (k == meshInst.size()) {
      DevInstance id;
      std::memset(&id, 0, sizeof id);
      id.o2w[0] = id.o2w[5] = id.o2w[10] = 1.f;
      id.w2o[0] = id.w2o[5] = id.w2o[10] = 1.f;
      id.vbase = in.vbase, id.ibase = in.ibase, id.ntri = in.ntri, id.light = -1;
      meshInst.push_back(id);
    }
    instBlas[i] = k;
  }
  const uint32_t nb = uint32_t(meshInst.size());
  std::vector<DevBlas> tab(nb);
  std::vector<HostBox> meshBox(nb);
  uint64_t shadeTotal = 0, nodes4 = 0, nodes2 = 0, tris = 0;
  uint32_t depth4 = 0;
  double ms = 0;
  for (uint32_t k = 0; k < nb; ++k) tab[k].shadeBase = uint32_t(shadeTotal), shadeTotal += meshInst[k].ntri;
  float4* shade = nullptr;
  HIPC(c, hipMalloc(&shade, std::max<uint64_t>(shadeTotal, 1) * 6 * sizeof(float4)));
  c->shadeTris = shade;
  struct Staging {  // the identity instance and its triangle-count prefix, freed on every return
    DBuf<DevInstance> one;
    DBuf<uint32_t> offs;
    ~Staging() { one.release(), offs.release(); }
  } tmp;
  DBuf<DevInstance>& one = tmp.one;
  DBuf<uint32_t>& offs = tmp.offs;
  MALLOC(c, one, 1);
  MALLOC(c, offs, 2);
  for (uint32_t k = 0; k < nb; ++k) {
    const uint32_t o2[2] = {0, meshInst[k].ntri};
    HIPC(c, hipMemcpyAsync(one.p, &meshInst[k], sizeof(DevInstance), hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(offs.p, o2, sizeof o2, hipMemcpyHostToDevice, s));
    BuildArgs a = tmpl;
    a.inst = one.p, a.numInst = 1, a.instTriOffset = offs.p, a.numTris = meshInst[k].ntri;
    a.branching = 4, a.quantize = 0;
    BuildResult r;
    char err[512] = {0};
    int rc = buildBvh(a, s, r, err, sizeof err);  // synchronises the stream before returning
    float4 root[4];
    if (rc == ASH_OK && hipMemcpy(root, r.nodes, sizeof root, hipMemcpyDeviceToHost) != hipSuccess) rc = ASH_E_HIP;
    if (rc == ASH_OK && hipMemcpy(shade + size_t(tab[k].shadeBase) * 6, r.shade,
                                  size_t(meshInst[k].ntri) * 6 * sizeof(float4), hipMemcpyDeviceToDevice) != hipSuccess)
      rc = ASH_E_HIP;
    if (r.nodes) (void)hipFree(r.nodes);
    if (r.nodesQ) (void)hipFree(r.nodesQ);
    if (r.shade) (void)hipFree(r.shade);
    if (r.nodes4) c->blasArrays.push_back(r.nodes4);
    if (r.tris) c->blasArrays.push_back(r.tris);
    if (rc) return fail(c, rc, "BLAS %u: %s", k, err[0] ? err : "copy failed");
    tab[k].nodes4 = r.nodes4, tab[k].tris = r.tris, tab[k].ntri = meshInst[k].ntri;
    // the root's box: union of its non-empty children (binary node layout)
    HostBox b{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    const float cl[2][6] = {{root[0].x, root[0].y, root[0].z, root[0].w, root[2].x, root[2].y},
                            {root[1].x, root[1].y, root[1].z, root[1].w, root[2].z, root[2].w}};
    const int32_t refs[2] = {bitsRef(root[3].x), bitsRef(root[3].y)};
    for (int ch = 0; ch < 2; ++ch) {
      if (refs[ch] == kEmptyRef) continue;
      for (int a = 0; a < 3; ++a)
        b.lo[a] = std::min(b.lo[a], double(cl[ch][2 * a])), b.hi[a] = std::max(b.hi[a], double(cl[ch][2 * a + 1]));
    }
    meshBox[k] = b;
    nodes4 += r.numNodes4, nodes2 += r.numNodes, tris += meshInst[k].ntri;
    depth4 = std::max(depth4, r.depth4);
    ms += r.ms;
  }
  // instance world boxes: the 8 corners of the mesh box through o2w, in
  // double, padded outwards by 1e-5 of their size and magnitude
  std::vector<HostBox> ibox(c->numInst);
  std::vector<uint32_t> items;
  for (uint32_t i = 0; i < c->numInst; ++i) {
    if (instBlas[i] == 0xffffffffu) continue;
    const HostBox& mb = meshBox[instBlas[i]];
    if (!(mb.lo[0] <= mb.hi[0])) continue;
    const float* m = c->hostInst[i].o2w;
    HostBox w{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    for (int corner = 0; corner < 8; ++corner) {
      const double p[3] = {corner & 1 ? mb.hi[0] : mb.lo[0], corner & 2 ? mb.hi[1] : mb.lo[1],
                           corner & 4 ? mb.hi[2] : mb.lo[2]};
      for (int a = 0; a < 3; ++a) {
        const double v = double(m[4 * a]) * p[0] + double(m[4 * a + 1]) * p[1] + double(m[4 * a + 2]) * p[2] +
                         double(m[4 * a + 3]);
        w.lo[a] = std::min(w.lo[a], v), w.hi[a] = std::max(w.hi[a], v);
      }
    }
    double mag = 0, ext = 0;
    for (int a = 0; a < 3; ++a)
      mag = std::max({mag, std::fabs(w.lo[a]), std::fabs(w.hi[a])}), ext = std::max(ext, w.hi[a] - w.lo[a]);
    const double pad = 1e-5 * (mag + ext) + 1e-30;
    for (int a = 0; a < 3; ++a) w.lo[a] -= pad, w.hi[a] += pad;
    ibox[i] = w;
    items.push_back(i);
  }
  std::vector<float4> tnodes;
  std::vector<uint32_t> order;
  uint32_t tdepth = 1;
  const HostBox far{{3.0e38, 3.0e38, 3.0e38}, {3.0e38, 3.0e38, 3.0e38}};
  if (items.size() <= 1) {
    tnodes.resize(4);
    if (items.empty()) {
      tlasNode(tnodes, 0, far, kEmptyRef, far, kEmptyRef);
    } else {
      order.push_back(items[0]);
      tlasNode(tnodes, 0, ibox[items[0]], encodeLeaf(0, 1), far, kEmptyRef);
    }
  } else {
    HostBox all;
    tdepth = 0;
    (void)tlasBuild(tnodes, order, ibox, items, 0, items.size(), all, 0, tdepth);
  }
  MALLOC(c, c->tlas, tnodes.size());
  MALLOC(c, c->tlasInst, std::max<size_t>(order.size(), 1));
  MALLOC(c, c->instBlas, std::max<size_t>(instBlas.size(), 1));
  MALLOC(c, c->blasTab, std::max<size_t>(tab.size(), 1));
  HIPC(c, hipMemcpy(c->tlas.p, tnodes.data(), tnodes.size() * sizeof(float4), hipMemcpyHostToDevice));
  if (!order.empty()) HIPC(c, hipMemcpy(c->tlasInst.p, order.data(), order.size() * 4, hipMemcpyHostToDevice));
  if (!instBlas.empty()) HIPC(c, hipMemcpy(c->instBlas.p, instBlas.data(), instBlas.size() * 4, hipMemcpyHostToDevice));
  if (!tab.empty()) HIPC(c, hipMemcpy(c->blasTab.p, tab.data(), tab.size() * sizeof(DevBlas), hipMemcpyHostToDevice));
  c->twoLevel = true;
  c->nodeCount = nodes4;
  c->maxDepth = tdepth + depth4;
  c->stackNeed = tdepth + 1 + 3 * depth4;  // one TLAS entry per level, three per 4-wide BLAS level
  if (info) {
    std::memset(info, 0, sizeof *info);
    info->num_triangles = c->numTris;
    info->num_nodes = nodes4 + tnodes.size() / 4;
    info->max_depth = c->maxDepth;
    info->node_bytes = 128;
    info->build_ms = ms;
    info->sah_cost = 0;
    info->device_bytes = nodes4 * 128 + tris * (48 + 96) + tnodes.size() * 16 + order.size() * 4 +
                         instBlas.size() * 4 + tab.size() * sizeof(DevBlas) + c->verts.n * 16 + c->idx.n * 4;
    info->num_wide_nodes = nodes4;
    info->wide_depth = depth4;
    info->branching = 4;
    info->blas_count = nb;
    info->blas_triangles = tris;
  }
  return ASH_OK;
}

// ------------------------------------------------- allocation-free render
// ash_render_noalloc keeps its wavefront (path state, queues, counters, the
// batch's work unit and shot record, traversal stack spill) in one fixed
// array per GPU instead of context buffers.  Host code never passes the
// array's name to a kernel: it names a host-side stand-in.  The device
// address comes from hipGetSymbolAddress, once per device with that device
// current (ash_create), and only that address reaches kernels.
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
constexpr uint32_t kNoAllocPaths = 1u << 18;  // paths in flight per batch
constexpr uint32_t kNoAllocBlocksPerCU = 2;   // persistent traversal blocks per CU
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
static std::mutex gScratchLookupMu;
static unsigned char* gScratchAddr[kMaxScratchDevices];
// Every call on one GPU shares its array; calls are serialised per device.
static std::mutex gScratchUseMu[kMaxScratchDevices];

static hipError_t scratchAddress(int device, unsigned char** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(gScratchLookupMu);
  if (!gScratchAddr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    gScratchAddr[device] = static_cast<unsigned char*>(p);
  }
  *out = gScratchAddr[device];
  return hipSuccess;
}

__global__ void kNoAllocReset(uint32_t* ctr, unsigned long long* cnt64) {
  for (uint32_t i = threadIdx.x; i < uint32_t(kNumCtr); i += blockDim.x) ctr[i] = 0;
  if (threadIdx.x < uint32_t(kNumCnt64)) cnt64[threadIdx.x] = 0;
}

// Stores the batch's work unit and shot record (kernel arguments) into the
// scratch, in stream order after the previous batch's readers.
__global__ void kNoAllocPut(WorkUnit* unit, WorkUnit u, ShotParams* shot, ShotParams sp) {
  if (threadIdx.x == 0) {
    *unit = u;
    *shot = sp;
  }
}

// Carves the scratch: per-path arrays for kNoAllocPaths paths, counters,
// one work unit, one shot record, then the traversal spill of `threads`
// persistent threads with `cap` entries each.  Returns the bytes used.
static size_t noAllocLayout(unsigned char* base, uint32_t threads, uint32_t cap, Wavefront& w, WorkUnit** unit,
                            ShotParams** shot) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += (bytes + 255) & ~size_t(255);
    return p;
  };
  const size_t n = kNoAllocPaths;
  std::memset(&w, 0, sizeof w);
  w.capacity = kNoAllocPaths;
  for (int q = 0; q < 2; ++q) {
    w.rayO[q] = reinterpret_cast<float4*>(take(n * sizeof(float4)));
    w.rayD[q] = reinterpret_cast<float4*>(take(n * sizeof(float4)));
  }
  for (int q = 0; q < 2; ++q) w.thr[q] = reinterpret_cast<float4*>(take(n * sizeof(float4)));
  w.rad = reinterpret_cast<float4*>(take(n * sizeof(float4)));
  w.ld = reinterpret_cast<float4*>(take(n * sizeof(float4)));
  w.hitF = reinterpret_cast<float4*>(take(n * sizeof(float4)));
  w.shO = reinterpret_cast<float4*>(take(n * sizeof(float4)));
  w.shD = reinterpret_cast<float4*>(take(n * sizeof(float4)));
  for (int q = 0; q < 2; ++q) w.info[q] = reinterpret_cast<uint4*>(take(n * sizeof(uint4)));
  w.hitI = reinterpret_cast<int4*>(take(n * sizeof(int4)));
  w.side = reinterpret_cast<int*>(take(n * sizeof(int)));
  w.ctr = reinterpret_cast<uint32_t*>(take(kNumCtr * sizeof(uint32_t)));
  w.cnt64 = reinterpret_cast<unsigned long long*>(take(kNumCnt64 * sizeof(unsigned long long)));
  *unit = reinterpret_cast<WorkUnit*>(take(sizeof(WorkUnit)));
  *shot = reinterpret_cast<ShotParams*>(take(sizeof(ShotParams)));
  w.spill = reinterpret_cast<int*>(take(size_t(threads) * cap * sizeof(int)));
  w.spillT = reinterpret_cast<float*>(take(size_t(threads) * cap * sizeof(float)));
  w.spillThreads = threads;
  w.spillCap = cap;
  return off;
}

extern "C" {

int ash_create(int device, ash_ctx** out) {
  if (!out) return ASH_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return ASH_E_HIP;
  if (device < 0 || device >= n) return ASH_E_INVALID;
  ash_ctx* c = new ash_ctx;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return ASH_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->numCU = prop.multiProcessorCount;
  c->scratchErr = scratchAddress(device, &c->scratch);
  *out = c;
  return ASH_OK;
}

void ash_destroy(ash_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  delete c;
}

const char* ash_last_error(const ash_ctx* c) { return c ? c->err.c_str() : "null context"; }

void ash_sun_and_sky_defaults(GpuSunAndSky* o) {
  if (!o) return;
  // Scene::init, scene.cpp:112-129
  *o = GpuSunAndSky{{1.f, 1.f, 1.f}, 0.0000101320f, 0.0f, 0.0f, 1.0f, 0.0f, {0.4f, 0.4f, 0.4f}, 0.1f,
                    {0.0f, 0.0f, 0.01f}, 0.8f, {0.00f, 0.78f, 0.62f}, 5.0f, 1.0f, 1, 1, 0};
}

// Writes the record and derives its direction-independent terms on the
// device (stream-ordered after the context's pending work).
static int setSky(ash_ctx* c, const GpuSunAndSky& ss) {
  if (!c->sky.p) MALLOC(c, c->sky, 1);
  DevSky k;
  std::memset(&k, 0, sizeof k);
  k.ss = ss;
  HIPC(c, hipMemcpyAsync(c->sky.p, &k, sizeof k, hipMemcpyHostToDevice, c->stream));
  ashd::launchSkyPrepare(c->sky.p, c->stream);
  HIPC(c, hipGetLastError());
  HIPC(c, hipStreamSynchronize(c->stream));
  return ASH_OK;
}

int ash_set_sun_and_sky(ash_ctx* c, const GpuSunAndSky* sky) {
  if (!c || !sky) return ASH_E_INVALID;
  if (!c->hasScene) return fail(c, ASH_E_STATE, "no scene uploaded");
  HIPC(c, hipSetDevice(c->device));
  return setSky(c, *sky);
}

int ash_upload_scene(ash_ctx* c, const ash_scene_desc* d) {
  if (!c || !d) return ASH_E_INVALID;
  HIPC(c, hipSetDevice(c->device));
  c->hasScene = c->hasAccel = false;
  if (d->num_materials == 0 || d->num_lights == 0) return fail(c, ASH_E_INVALID, "material 0 and light 0 (dummies) are required");
  if (d->num_materials > kShadeIdMask || d->num_lights > kShadeIdMask)
    return fail(c, ASH_E_UNSUPPORTED, "more than 2^24 - 1 materials or lights");
  // flatten meshes
  std::vector<uint32_t> vbase(d->num_meshes), ibase(d->num_meshes);
  size_t nv = 0, ni = 0;
  for (uint32_t m = 0; m < d->num_meshes; ++m) {
    const ash_mesh_desc& md = d->meshes[m];
    if (md.num_indices % 3) return fail(c, ASH_E_INVALID, "mesh %u: index count not a multiple of 3", m);
    for (uint32_t k = 0; k < md.num_indices; ++k)
      if (md.indices[k] >= md.num_vertices) return fail(c, ASH_E_INVALID, "mesh %u: index out of range", m);
    vbase[m] = uint32_t(nv);
    ibase[m] = uint32_t(ni);
    nv += md.num_vertices;
    ni += md.num_indices;
  }
  if (nv >= 0xffffffffull || ni >= 0xffffffffull) return fail(c, ASH_E_UNSUPPORTED, "scene too large");
  std::vector<float4> hv(nv * 2);
  std::vector<uint32_t> hi(ni);
  for (uint32_t m = 0; m < d->num_meshes; ++m) {
    const ash_mesh_desc& md = d->meshes[m];
    for (uint32_t k = 0; k < md.num_vertices; ++k) {
      const GpuVertex& v = md.vertices[k];
      hv[2 * (vbase[m] + k)] = make_float4(v.pos[0], v.pos[1], v.pos[2], v.uv[0]);
      hv[2 * (vbase[m] + k) + 1] = make_float4(v.normal[0], v.normal[1], v.normal[2], v.uv[1]);
    }
    std::memcpy(&hi[ibase[m]], md.indices, md.num_indices * sizeof(uint32_t));
  }
  std::vector<DevInstance> hinst(d->num_instances);
  std::vector<uint32_t> off(d->num_instances + 1, 0);
  uint64_t tri = 0;
  for (uint32_t i = 0; i < d->num_instances; ++i) {
    const ash_instance_desc& id = d->instances[i];
    if (id.mesh_id >= d->num_meshes) return fail(c, ASH_E_INVALID, "instance %u: mesh id out of range", i);
    if (id.light_id >= int32_t(d->num_lights)) return fail(c, ASH_E_INVALID, "instance %u: light id out of range", i);
    if (id.light_id < 0) {
      if (id.material_id >= d->num_materials) return fail(c, ASH_E_INVALID, "instance %u: material id out of range", i);
      uint32_t t = d->materials[id.material_id].type;
      if (t >= MaterialTypeNum)
        return fail(c, ASH_E_UNSUPPORTED, "instance %u: material type %u is outside the supported plugin set", i, t);
    }
    DevInstance& di = hinst[i];
    std::memset(&di, 0, sizeof di);
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 4; ++cc) di.o2w[r * 4 + cc] = id.transform[cc * 4 + r];
    invert3x4(di.o2w, di.w2o);
    di.vbase = vbase[id.mesh_id];
    di.ibase = ibase[id.mesh_id];
    di.material = id.light_id < 0 ? id.material_id : 0;
    di.light = id.light_id;
    di.ntri = d->meshes[id.mesh_id].num_indices / 3;
    if (id.light_id >= 0) di.shade = kShadeLight | uint32_t(id.light_id);
    else {
      const uint32_t t = d->materials[id.material_id].type;
      di.shade = id.material_id | ((t & kShadeTypeMask) << kShadeTypeShift) | (isGroup2Type(t) ? kShadeSide : 0u);
    }
    di.triBase = uint32_t(tri);
    off[i] = uint32_t(tri);
    tri += di.ntri;
  }
  if (tri >= (1ull << 28)) return fail(c, ASH_E_UNSUPPORTED, "more than 2^28 triangles");
  off[d->num_instances] = uint32_t(tri);
  for (uint32_t t = 0; t < d->num_textures; ++t)
    if (!d->textures[t].rgba || !d->textures[t].width || !d->textures[t].height)
      return fail(c, ASH_E_INVALID, "texture %u has no pixels", t);
  for (uint32_t m = 0; m < d->num_materials; ++m) {
    const GpuMaterial& g = d->materials[m];
    const int32_t ids[] = {g.diffuseTextureId, g.roughnessTextureId, g.metalnessTextureId, g.radianceTextureId,
                           g.normalTextureId, g.opacityTextureId};
    for (int32_t t : ids)
      if (t >= int32_t(d->num_textures)) return fail(c, ASH_E_INVALID, "material %u: texture id %d out of range", m, t);
  }
  hipStream_t s = c->stream;
  MALLOC(c, c->verts, hv.size());
  MALLOC(c, c->idx, hi.size());
  MALLOC(c, c->inst, hinst.size());
  MALLOC(c, c->instTriOffset, off.size());
  MALLOC(c, c->mats, d->num_materials);
  MALLOC(c, c->lights, d->num_lights);
  HIPC(c, hipMemcpyAsync(c->verts.p, hv.data(), hv.size() * sizeof(float4), hipMemcpyHostToDevice, s));
  HIPC(c, hipMemcpyAsync(c->idx.p, hi.data(), hi.size() * 4, hipMemcpyHostToDevice, s));
  HIPC(c, hipMemcpyAsync(c->inst.p, hinst.data(), hinst.size() * sizeof(DevInstance), hipMemcpyHostToDevice, s));
  HIPC(c, hipMemcpyAsync(c->instTriOffset.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, s));
  c->hostInst = hinst;
  std::vector<DevMaterial> hmats(d->num_materials);
  for (uint32_t k = 0; k < d->num_materials; ++k) {
    std::memset(&hmats[k], 0, sizeof(DevMaterial));
    hmats[k].m = d->materials[k];
  }
  HIPC(c, hipMemcpyAsync(c->mats.p, hmats.data(), hmats.size() * sizeof(DevMaterial), hipMemcpyHostToDevice, s));
  HIPC(c, hipMemcpyAsync(c->lights.p, d->lights, d->num_lights * sizeof(GpuLight), hipMemcpyHostToDevice, s));
  for (auto& t : c->texData) t.release();
  c->texData.assign(d->num_textures, DBuf<float4>());
  std::vector<DevImage> tab(d->num_textures);
  for (uint32_t t = 0; t < d->num_textures; ++t) {
    const ash_image_desc& im = d->textures[t];
    size_t npx = size_t(im.width) * im.height;
    MALLOC(c, c->texData[t], npx);
    HIPC(c, hipMemcpyAsync(c->texData[t].p, im.rgba, npx * sizeof(float4), hipMemcpyHostToDevice, s));
    tab[t].px = c->texData[t].p;
    tab[t].w = im.width;
    tab[t].h = im.height;
  }
  MALLOC(c, c->texTab, tab.size());
  if (!tab.empty())
    HIPC(c, hipMemcpyAsync(c->texTab.p, tab.data(), tab.size() * sizeof(DevImage), hipMemcpyHostToDevice, s));
  // env map + importance tables (EnvMap::EnvMap, texture.cpp:144-226)
  c->hasEnv = d->envmap.rgba != nullptr;
  std::vector<float> marg, cond;
  if (c->hasEnv) {
    c->envW = d->envmap.width;
    c->envH = d->envmap.height;
    if (!c->envW || !c->envH) return fail(c, ASH_E_INVALID, "env map has zero size");
    ash::buildEnvTables(c->envW, c->envH, d->envmap.rgba, marg, cond);
    const uint32_t W = c->envW, H = c->envH;
    std::vector<float4> px(envTexels(W, H), make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<float> cx(condTexels(W, H), 0.f);
    for (uint32_t y = 0; y < H; ++y)
      for (uint32_t x = 0; x < W; ++x) {
        const size_t k = size_t(y) * W + x;
        px[envTexel(W, x, y)] =
            make_float4(d->envmap.rgba[4 * k], d->envmap.rgba[4 * k + 1], d->envmap.rgba[4 * k + 2], cond[2 * k + 1]);
        cx[condTexel(W, x, y)] = cond[2 * k];
      }
    MALLOC(c, c->env, px.size());
    MALLOC(c, c->marg, H);
    MALLOC(c, c->condX, cx.size());
    HIPC(c, hipMemcpy(c->env.p, px.data(), px.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->marg.p, marg.data(), marg.size() * 4, hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->condX.p, cx.data(), cx.size() * 4, hipMemcpyHostToDevice));
  } else {
    c->envW = c->envH = 1;
    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 z2 = make_float2(0.f, 0.f);
    float z1 = 0.f;
    MALLOC(c, c->env, 1);
    MALLOC(c, c->marg, 1);
    MALLOC(c, c->condX, 1);
    HIPC(c, hipMemcpy(c->env.p, &z, sizeof z, hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->marg.p, &z2, sizeof z2, hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->condX.p, &z1, sizeof z1, hipMemcpyHostToDevice));
  }
  {  // every scene starts from the default sun & sky (in_use 0)
    GpuSunAndSky ss;
    ash_sun_and_sky_defaults(&ss);
    int rc = setSky(c, ss);
    if (rc != ASH_OK) return rc;
  }
  HIPC(c, hipStreamSynchronize(s));
  c->numTris = uint32_t(tri);
  c->numInst = d->num_instances;
  c->hasScene = true;
  return ASH_OK;
}

int ash_build_accel(ash_ctx* c, const ash_build_opts* o, ash_build_info* info) {
  if (!c) return ASH_E_INVALID;
  if (!c->hasScene) return fail(c, ASH_E_STATE, "build_accel before upload_scene");
  HIPC(c, hipSetDevice(c->device));
  if (c->nodes) (void)hipFree(c->nodes), c->nodes = nullptr;
  if (c->nodes4) (void)hipFree(c->nodes4), c->nodes4 = nullptr;
  if (c->nodesQ) (void)hipFree(c->nodesQ), c->nodesQ = nullptr;
  if (c->tris) (void)hipFree(c->tris), c->tris = nullptr;
  if (c->shadeTris) (void)hipFree(c->shadeTris), c->shadeTris = nullptr;
  c->releaseTwoLevel();
  c->hasAccel = false;
  BuildArgs a;
  a.inst = c->inst.p;
  a.numInst = c->numInst;
  a.instTriOffset = c->instTriOffset.p;
  a.vertices = c->verts.p;
  a.indices = c->idx.p;
  a.numTris = c->numTris;
  a.leafMax = o && o->leaf_max_tris ? o->leaf_max_tris : 4;
  a.quality = o ? o->quality : 2;
  if (a.quality > 2) return fail(c, ASH_E_INVALID, "quality must be 0, 1 or 2");
  a.branching = o && o->branching ? o->branching : 4;
  if (a.branching != 2 && a.branching != 4) return fail(c, ASH_E_INVALID, "branching must be 2 or 4");
  if (c->numTris >= (1u << 28)) return fail(c, ASH_E_UNSUPPORTED, "more than 2^28 - 1 triangles");
  uint32_t fmt = o ? o->node_format : 0;
  if (fmt > 2) return fail(c, ASH_E_INVALID, "node_format must be 0, 1 or 2");
  a.quantize = a.branching == 4 && fmt == 2;
  if (o && o->two_level) {
    if (a.branching != 4 || fmt == 2) return fail(c, ASH_E_INVALID, "two_level needs the float 4-wide tree");
    const int rc = buildTwoLevel(c, a, info);
    if (rc) {
      c->releaseTwoLevel();
      if (c->shadeTris) (void)hipFree(c->shadeTris), c->shadeTris = nullptr;
      return rc;
    }
    c->hasAccel = true;
    return ASH_OK;
  }
  BuildResult r;
  char err[512] = {0};
  int rc = buildBvh(a, c->stream, r, err, sizeof err);
  if (rc) {
    if (r.nodes) (void)hipFree(r.nodes);
    if (r.nodes4) (void)hipFree(r.nodes4);
    if (r.nodesQ) (void)hipFree(r.nodesQ);
    if (r.tris) (void)hipFree(r.tris);
    if (r.shade) (void)hipFree(r.shade);
    return fail(c, rc, "%s", err);
  }
  c->nodes = r.nodes;
  c->nodes4 = r.nodes4;
  c->nodesQ = r.nodesQ;
  c->tris = r.tris;
  c->shadeTris = r.shade;
  c->nodeCount = r.numNodes;
  c->maxDepth = r.maxDepth;
  // A binary walk stacks at most one entry per level, a 4-wide walk three.
  c->stackNeed = r.nodes4 ? 3 * r.depth4 : r.maxDepth;
  c->hasAccel = true;
  if (info) {
    info->num_triangles = c->numTris;
    info->num_nodes = r.numNodes;
    info->max_depth = r.maxDepth;
    info->node_bytes = r.nodesQ ? 64u : (r.nodes4 ? 128u : 64u);
    info->build_ms = r.ms;
    info->sah_cost = r.sahCost;
    info->device_bytes = uint64_t(r.numNodes) * 64 + uint64_t(r.numNodes4) * (r.nodesQ ? 192 : 128) +
                         uint64_t(c->numTris) * (48 + 96) +
                         c->verts.n * 16 + c->idx.n * 4;
    info->num_wide_nodes = r.numNodes4;
    info->wide_depth = r.depth4;
    info->branching = a.branching;
    info->blas_count = 0;
    info->blas_triangles = 0;
  }
  return ASH_OK;
}

int ash_set_film(ash_ctx* c, uint32_t w, uint32_t h) {
  if (!c || !w || !h) return ASH_E_INVALID;
  HIPC(c, hipSetDevice(c->device));
  size_t n = size_t(w) * h * ASH_NUM_OUTPUT_IMAGES;
  if (c->width != w || c->height != h || !c->images.p) MALLOC(c, c->images, n);
  c->width = w;
  c->height = h;
  HIPC(c, hipMemsetAsync(c->images.p, 0, n * sizeof(float4), c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return ASH_OK;
}

static ShotParams makeShot(const ash_ctx* c, const GpuCamera* cam, const GpuPushConstantRaytrace* pc,
                           const ash_render_opts* o, float4* images, uint32_t w, uint32_t h) {
  ShotParams sp;
  std::memset(&sp, 0, sizeof sp);
  sp.cam = *cam;
  sp.pc = *pc;
  sp.pc.spp = 1;
  sp.depthChannel = o ? o->depthOutChannel : -1;
  if (o == nullptr) sp.depthChannel = -1;
  sp.width = w;
  sp.height = h;
  sp.images = images;
  (void)c;
  return sp;
}

int ash_render(ash_ctx* c, const GpuCamera* cam, const GpuPushConstantRaytrace* pc, int32_t first, int32_t n,
               const ash_render_opts* o, ash_stats* st) {
  if (!c || !cam || !pc) return ASH_E_INVALID;
  if (!c->images.p) return fail(c, ASH_E_STATE, "render before set_film");
  HIPC(c, hipSetDevice(c->device));
  if (o && o->depthOutChannel >= int32_t(ASH_NUM_OUTPUT_IMAGES - 1)) return fail(c, ASH_E_INVALID, "depth channel out of range");
  ShotParams sp = makeShot(c, cam, pc, o, c->images.p, c->width, c->height);
  return renderCommon(c, 1, &sp, first, n, o, st);
}

int ash_render_noalloc(ash_ctx* c, const GpuCamera* cam, const GpuPushConstantRaytrace* pc, int32_t first, int32_t n,
                       uint32_t w, uint32_t h, void* images, ash_stats* st) {
  if (!c || !cam || !pc || !images || !w || !h) return ASH_E_INVALID;
  if (!c->hasScene || !c->hasAccel) return fail(c, ASH_E_STATE, "render called before upload_scene/build_accel");
  if (n <= 0 || first < 0 || int64_t(first) + n > int64_t(INT32_MAX)) return fail(c, ASH_E_INVALID, "bad frame range");
  if (uint64_t(w) * h > (uint64_t(1) << 30)) return fail(c, ASH_E_INVALID, "film above 2^30 pixels");
  if (!c->scratch)
    return fail(c, ASH_E_HIP, "allocation-free scratch unavailable on device %d: %s", c->device,
                hipGetErrorString(c->scratchErr));
  if (int rc = checkPc(c, pc)) return rc;
  const uint32_t grid = uint32_t(std::max(1, c->numCU)) * kNoAllocBlocksPerCU;
  const uint32_t cap =
      std::max<uint32_t>(uint32_t(kSpill), c->stackNeed > uint32_t(kStackLdsMin) ? c->stackNeed - kStackLdsMin : 0);
  Wavefront wf;
  WorkUnit* unitDev = nullptr;
  ShotParams* shotDev = nullptr;
  const size_t need = noAllocLayout(c->scratch, grid * kTraceBlock, cap, wf, &unitDev, &shotDev);
  if (need > kNoAllocScratchBytes)
    return fail(c, ASH_E_INVALID, "allocation-free render needs %zu bytes of scratch (%u traversal threads x %u stack "
                "entries), above its %zu-byte limit", need, grid * kTraceBlock, cap, kNoAllocScratchBytes);
  std::lock_guard<std::mutex> use(gScratchUseMu[c->device]);
  HIPC(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const ShotParams sp = makeShot(c, cam, pc, nullptr, static_cast<float4*>(images), w, h);
  const SceneView sc = viewOf(c);
  const uint32_t P = w * h, iters = uint32_t(std::max(pc->maxPathDepth, 1));
  const int32_t end = first + n;
  auto t0 = std::chrono::steady_clock::now();
  kNoAllocReset<<<1, 256, 0, s>>>(wf.ctr, wf.cnt64);
  uint64_t batches = 0;
  int rc = ASH_OK;
  // The batches of a single-shot render (planBatches with one shot): whole
  // frames grouped while a frame fits the queue, else runs of pixels.
  auto batch = [&](uint32_t pix0, uint32_t np, int32_t f0, uint32_t nf) {
    if (rc != ASH_OK) return;
    WorkUnit u;
    std::memset(&u, 0, sizeof u);
    u.shot = 0, u.pix0 = pix0, u.np = np, u.f0 = f0, u.nf = nf;
    kNoAllocPut<<<1, 64, 0, s>>>(unitDev, u, shotDev, sp);
    BatchDesc b;
    std::memset(&b, 0, sizeof b);
    b.n = np * nf;
    b.nAcc = np;
    b.nUnits = 1;
    b.shots = 1;
    b.units = unitDev;
    b.shotTab = shotDev;
    b.one = sp;
    launchRaygen(b, wf, s);
    rc = launchLoop(c, wf, sc, b, iters, grid, uint32_t(kStackLds), uint32_t(kSchemeXcdSlices), b.n, false, nullptr);
    launchAccumulate(b, wf, s);
    ++batches;
  };
  if (P <= kNoAllocPaths) {
    const int32_t F = int32_t(std::min<uint32_t>(uint32_t(n), kNoAllocPaths / P));
    for (int32_t f0 = first; f0 < end; f0 += F) batch(0, P, f0, uint32_t(std::min(F, end - f0)));
  } else {
    for (int32_t f = first; f < end; ++f)
      for (uint32_t p0 = 0; p0 < P; p0 += kNoAllocPaths) batch(p0, std::min(kNoAllocPaths, P - p0), f, 1);
  }
  if (rc) return fail(c, rc, "allocation-free render: launch failure");
  HIPC(c, hipGetLastError());
  HIPC(c, hipStreamSynchronize(s));
  if (st) {
    unsigned long long cnt[kNumCnt64];
    HIPC(c, hipMemcpy(cnt, wf.cnt64, sizeof cnt, hipMemcpyDeviceToHost));
    std::memset(st, 0, sizeof *st);
    st->render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    st->paths = uint64_t(P) * uint64_t(n);
    st->batches = batches;
    st->extension_rays = cnt[4];
    st->shadow_rays = cnt[5];
    st->dielectric_shadow_rays = cnt[6];
  }
  return ASH_OK;
}

int ash_render_shots(ash_ctx* c, uint32_t nShots, const GpuCamera* cams, const GpuPushConstantRaytrace* pcs,
                     int32_t nFrames, const ash_render_opts* o, ash_stats* st) {
  if (!c || !cams || !pcs || nShots == 0 || nShots > 0xffff) return ASH_E_INVALID;
  if (!c->width) return fail(c, ASH_E_STATE, "render_shots before set_film (film size)");
  if (o && o->depthOutChannel >= int32_t(ASH_NUM_OUTPUT_IMAGES - 1)) return fail(c, ASH_E_INVALID, "depth channel out of range");
  HIPC(c, hipSetDevice(c->device));
  size_t plane = size_t(c->width) * c->height;
  size_t n = plane * ASH_NUM_OUTPUT_IMAGES * nShots;
  if (c->shotCount != nShots || c->shotW != c->width || c->shotH != c->height || !c->shotImages.p)
    MALLOC(c, c->shotImages, n);
  c->shotCount = nShots;
  c->shotW = c->width;
  c->shotH = c->height;
  HIPC(c, hipMemsetAsync(c->shotImages.p, 0, n * sizeof(float4), c->stream));
  std::vector<ShotParams> sps(nShots);
  for (uint32_t s = 0; s < nShots; ++s)
    sps[s] = makeShot(c, &cams[s], &pcs[s], o, c->shotImages.p + s * plane * ASH_NUM_OUTPUT_IMAGES, c->width, c->height);
  return renderCommon(c, nShots, sps.data(), 0, nFrames, o, st);
}

int ash_autotune(ash_ctx* c, uint32_t nShots, const GpuCamera* cams, const GpuPushConstantRaytrace* pcs,
                 int32_t probeFrames, ash_render_opts* io, ash_tune_result* res) {
  if (!c || !cams || !pcs || !io || nShots == 0 || probeFrames <= 0) return ASH_E_INVALID;
  ash_render_opts o = *io;
  o.timing = 0;
  o.count_traversal = 0;
  if (o.traversal > uint32_t(kSchemeRefill)) return fail(c, ASH_E_INVALID, "traversal must be 0, 1 or 2");
  double best = 1e300, worst = 0;
  uint32_t bq = 0, bb = 0, n = 0;
  uint32_t bl = 0, bs = o.traversal;
  uint64_t batches = 0;
  auto probe = [&](double& ms) -> int {
    ash_stats st;
    for (int rep = 0; rep < 2; ++rep) {
      int rc = nShots == 1 ? ash_render(c, cams, pcs, 0, probeFrames, &o, &st)
                           : ash_render_shots(c, nShots, cams, pcs, probeFrames, &o, &st);
      if (rc) return rc;
      ms = st.render_ms;
    }
    batches = st.batches;
    ++n;
    worst = std::max(worst, ms);
    return ASH_OK;
  };
  // queue sizes from 2^18 up; once the probe fits in one batch, larger
  // queues would launch the same work
  for (uint32_t lq = 18; lq <= 26; ++lq) {
    for (uint32_t bpc : {4u, 8u, 16u})
      for (uint32_t lds : {uint32_t(kStackLdsMin), uint32_t(kStackLds)}) {
        o.queue_paths = 1u << lq;
        o.blocks_per_cu = bpc;
        o.lds_stack = lds;
        double ms = 0;
        if (int rc = probe(ms)) return rc;
        if (ms < best) best = ms, bq = o.queue_paths, bb = bpc, bl = lds;
      }
    if (batches <= 1) break;
  }
  o.queue_paths = bq, o.blocks_per_cu = bb, o.lds_stack = bl;
  const uint32_t first = o.traversal;
  for (uint32_t sch = 0; sch <= uint32_t(kSchemeRefill); ++sch) {
    if (sch == first) continue;
    o.traversal = sch;
    double ms = 0;
    if (int rc = probe(ms)) return rc;
    if (ms < best) best = ms, bs = sch;
  }
  io->queue_paths = bq;
  io->blocks_per_cu = bb;
  io->lds_stack = bl;
  io->traversal = bs;
  if (res) {
    res->queue_paths = bq;
    res->blocks_per_cu = bb;
    res->candidates = n;
    res->lds_stack = bl;
    res->traversal = bs;
    res->pad = 0;
    res->best_ms = best;
    res->worst_ms = worst;
  }
  return ASH_OK;
}

int ash_read_shot_output(ash_ctx* c, uint32_t shot, int32_t image, float* out) {
  if (!c || !out || image < 0 || image >= ASH_NUM_OUTPUT_IMAGES || shot >= c->shotCount) return ASH_E_INVALID;
  HIPC(c, hipSetDevice(c->device));
  size_t plane = size_t(c->shotW) * c->shotH;
  HIPC(c, hipMemcpy(out, c->shotImages.p + (shot * ASH_NUM_OUTPUT_IMAGES + size_t(image)) * plane, plane * sizeof(float4),
                    hipMemcpyDeviceToHost));
  return ASH_OK;
}

int ash_read_output(ash_ctx* c, int32_t image, float* out) {
  if (!c || !out || image < 0 || image >= ASH_NUM_OUTPUT_IMAGES) return ASH_E_INVALID;
  if (!c->images.p) return fail(c, ASH_E_STATE, "no film");
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipStreamSynchronize(c->stream));
  size_t plane = size_t(c->width) * c->height;
  HIPC(c, hipMemcpy(out, c->images.p + size_t(image) * plane, plane * sizeof(float4), hipMemcpyDeviceToHost));
  return ASH_OK;
}

int ash_output_device_ptr(ash_ctx* c, int32_t image, void** p) {
  if (!c || !p || image < 0 || image >= ASH_NUM_OUTPUT_IMAGES) return ASH_E_INVALID;
  if (!c->images.p) return fail(c, ASH_E_STATE, "no film");
  *p = c->images.p + size_t(image) * c->width * c->height;
  return ASH_OK;
}

int ash_first_hit(ash_ctx* c, const GpuCamera* cam, ash_hit* out) {
  if (!c || !cam || !out) return ASH_E_INVALID;
  if (!c->hasAccel) return fail(c, ASH_E_STATE, "first_hit before build_accel");
  if (!c->width) return fail(c, ASH_E_STATE, "first_hit before set_film");
  HIPC(c, hipSetDevice(c->device));
  uint32_t P = c->width * c->height;
  int rc = ensureWavefront(c, P);
  if (rc) return rc;
  uint32_t grid = uint32_t(std::max(1, c->numCU)) * 8;
  rc = ensureSpill(c, grid);
  if (rc) return rc;
  GpuPushConstantRaytrace pc;
  std::memset(&pc, 0, sizeof pc);
  pc.maxPathDepth = 1;
  ShotParams sp = makeShot(c, cam, &pc, nullptr, c->images.p, c->width, c->height);
  WorkUnit u;
  std::memset(&u, 0, sizeof u);
  u.np = P;
  u.nf = 1;
  if (c->shotTab.n < 1) MALLOC(c, c->shotTab, 1);
  if (c->unitBuf.n < 1) MALLOC(c, c->unitBuf, 1);
  HIPC(c, hipStreamSynchronize(c->stream));
  HIPC(c, hipMemcpy(c->shotTab.p, &sp, sizeof sp, hipMemcpyHostToDevice));
  HIPC(c, hipMemcpy(c->unitBuf.p, &u, sizeof u, hipMemcpyHostToDevice));
  DBuf<ash_hit> dout;
  MALLOC(c, dout, P);
  BatchDesc b;
  std::memset(&b, 0, sizeof b);
  b.n = P;
  b.nAcc = P;
  b.nUnits = 1;
  b.shots = 1;
  b.units = c->unitBuf.p;
  b.shotTab = c->shotTab.p;
  b.one = sp;
  SceneView sc = viewOf(c);
  launchRaygen(b, c->wf, c->stream);
  launchResetCounters(c->wf, 0, c->stream);
  launchExtendCamera(sc, c->wf, kStackLds, kSchemeXcdSlices, grid, false, c->numTris <= kCoarseFetchTris,
                     c->stream);
  launchFirstHit(sc, c->shotTab.p, c->wf, dout.p, P, c->stream);
  HIPC(c, hipGetLastError());
  HIPC(c, hipMemcpyAsync(out, dout.p, P * sizeof(ash_hit), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  dout.release();
  return ASH_OK;
}

int ash_env_tables(ash_ctx* c, float* m, float* cd) {
  if (!c || !m || !cd) return ASH_E_INVALID;
  if (!c->hasEnv) return fail(c, ASH_E_STATE, "no env map");
  HIPC(c, hipSetDevice(c->device));
  HIPC(c, hipMemcpy(m, c->marg.p, size_t(c->envH) * sizeof(float2), hipMemcpyDeviceToHost));
  // the conditional table as (inverse CDF, pdf) pairs, from its two device planes
  const uint32_t W = c->envW, H = c->envH;
  std::vector<float4> px(c->env.n);
  std::vector<float> cx(c->condX.n);
  HIPC(c, hipMemcpy(px.data(), c->env.p, px.size() * sizeof(float4), hipMemcpyDeviceToHost));
  HIPC(c, hipMemcpy(cx.data(), c->condX.p, cx.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (uint32_t y = 0; y < H; ++y)
    for (uint32_t x = 0; x < W; ++x) {
      const size_t k = size_t(y) * W + x;
      cd[2 * k] = cx[condTexel(W, x, y)];
      cd[2 * k + 1] = px[envTexel(W, x, y)].w;
    }
  return ASH_OK;
}

void ash_pod_layout(uint32_t* o) {
  o[0] = sizeof(GpuVertex), o[1] = offsetof(GpuVertex, tangent);
  o[2] = sizeof(GpuMaterial), o[3] = offsetof(GpuMaterial, type);
  o[4] = sizeof(GpuLight), o[5] = offsetof(GpuLight, doubleSide);
  o[6] = sizeof(GpuCamera), o[7] = offsetof(GpuCamera, padding);
  o[8] = sizeof(GpuPushConstantRaytrace), o[9] = offsetof(GpuPushConstantRaytrace, uvOutChannel);
}

const char* ash_version(void) { return "asuna-mi355x 0.2 (gfx950, wavefront path tracer, LBVH+SAH collapse, 4-wide BVH)"; }

}  // extern "C"

namespace ashd {

hipStream_t ctxStream(ash_ctx* c) { return c->stream; }
int ctxNumCU(ash_ctx* c) { return std::max(1, c->numCU); }
int ctxPrepare(ash_ctx* c, uint32_t capacity, uint32_t grid) {
  if (!c->hasScene || !c->hasAccel) return fail(c, ASH_E_STATE, "scene not uploaded/built");
  int rc = ensureWavefront(c, capacity);
  return rc ? rc : ensureSpill(c, grid);
}
int ctxCheckPc(ash_ctx* c, const GpuPushConstantRaytrace* pc) { return checkPc(c, pc); }
int ctxRunBatch(ash_ctx* c, const BatchDesc& b, int maxDepth, uint32_t grid, const ash_render_opts* o,
                ash_stats* st) {
  return runBatch(c, b, maxDepth, grid, o, st);
}
ShotParams ctxMakeShot(const GpuCamera* cam, const GpuPushConstantRaytrace* pc, int32_t depthChannel,
                       float4* images, uint32_t w, uint32_t h) {
  ash_render_opts o;
  std::memset(&o, 0, sizeof o);
  o.depthOutChannel = depthChannel;
  return makeShot(nullptr, cam, pc, &o, images, w, h);
}
const char* ctxError(ash_ctx* c) { return c->err.c_str(); }
void ctxFilm(ash_ctx* c, uint32_t* w, uint32_t* h) {
  *w = c->images.p ? c->width : 0;
  *h = c->images.p ? c->height : 0;
}

}  // namespace ashd

