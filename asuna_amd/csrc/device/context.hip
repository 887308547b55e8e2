// C ABI of the render library (include/asuna_hip.h): device context, scene
// upload, BVH build, wavefront render loop, output readback.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../host/scene.h"
#include "render.h"

using namespace ashd;

static_assert(sizeof(GpuVertex) == 44, "GpuVertex must be 44 bytes (vertex.h)");
static_assert(sizeof(GpuMaterial) == 132, "GpuMaterial must be 132 bytes (material.h)");
static_assert(sizeof(GpuLight) == 76, "GpuLight must be 76 bytes (light.h)");
static_assert(sizeof(GpuCamera) == 224, "GpuCamera must be 224 bytes (camera.h)");
static_assert(sizeof(GpuPushConstantRaytrace) == 84, "push constants must be 84 bytes (pushconstant.h)");
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
  DBuf<GpuMaterial> mats;
  DBuf<GpuLight> lights;
  std::vector<DBuf<float4>> texData;
  DBuf<DevImage> texTab;
  DBuf<float4> env;
  DBuf<float2> marg, cond;
  uint32_t envW = 1, envH = 1;
  bool hasEnv = false;
  uint32_t numTris = 0, numInst = 0;
  uint32_t maxDepth = 0;
  float4* nodes = nullptr;
  float4* tris = nullptr;
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
  DBuf<float4> wRo, wRd, wThr, wRad, wLd, wHitF, wShO, wShD;
  DBuf<uint4> wInfo;
  DBuf<int4> wHitI;
  DBuf<int> wQ0, wQ1, wSpill;
  DBuf<uint32_t> wCtr;
  DBuf<unsigned long long> wCnt64;
  uint32_t traceGrid = 0;

  ~ash_ctx() {
    if (nodes) (void)hipFree(nodes);
    if (tris) (void)hipFree(tris);
    for (auto& t : texData) t.release();
    for (auto* b : {&verts, &env, &images, &shotImages, &wRo, &wRd, &wThr, &wRad, &wLd, &wHitF, &wShO, &wShD})
      b->release();
    inst.release(), idx.release(), instTriOffset.release(), mats.release(), lights.release(), texTab.release();
    marg.release(), cond.release(), shotTab.release(), wInfo.release(), wHitI.release();
    wQ0.release(), wQ1.release(), wSpill.release(), wCtr.release(), wCnt64.release();
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
  v.nodes = c->nodes;
  v.tris = c->tris;
  v.inst = c->inst.p;
  v.verts = c->verts.p;
  v.idx = c->idx.p;
  v.mats = c->mats.p;
  v.lights = c->lights.p;
  v.textures = c->texTab.p;
  v.numTextures = uint32_t(c->texTab.n);
  v.hasEnv = c->hasEnv ? 1u : 0u;
  v.env.px = c->env.p;
  v.env.w = c->envW;
  v.env.h = c->envH;
  v.marg = c->marg.p;
  v.cond = c->cond.p;
  return v;
}

static int ensureWavefront(ash_ctx* c, uint32_t cap) {
  if (c->wf.capacity >= cap && c->wRo.p) return ASH_OK;
  MALLOC(c, c->wRo, cap);
  MALLOC(c, c->wRd, cap);
  MALLOC(c, c->wThr, cap);
  MALLOC(c, c->wRad, cap);
  MALLOC(c, c->wLd, cap);
  MALLOC(c, c->wHitF, cap);
  MALLOC(c, c->wShO, cap);
  MALLOC(c, c->wShD, cap);
  MALLOC(c, c->wInfo, cap);
  MALLOC(c, c->wHitI, cap);
  MALLOC(c, c->wQ0, cap);
  MALLOC(c, c->wQ1, cap);
  MALLOC(c, c->wCtr, 16);
  MALLOC(c, c->wCnt64, 4);
  c->traceGrid = uint32_t(std::max(1, c->numCU) * 8);
  uint32_t spillThreads = c->traceGrid * kTraceBlock;
  MALLOC(c, c->wSpill, size_t(spillThreads) * kSpill);
  Wavefront& w = c->wf;
  w.capacity = cap;
  w.ro = c->wRo.p;
  w.rd = c->wRd.p;
  w.thr = c->wThr.p;
  w.rad = c->wRad.p;
  w.ld = c->wLd.p;
  w.info = c->wInfo.p;
  w.hitF = c->wHitF.p;
  w.hitI = c->wHitI.p;
  w.queue[0] = c->wQ0.p;
  w.queue[1] = c->wQ1.p;
  w.shO = c->wShO.p;
  w.shD = c->wShD.p;
  w.ctr = c->wCtr.p;
  w.spill = c->wSpill.p;
  w.spillThreads = spillThreads;
  w.cnt64 = c->wCnt64.p;
  HIPC(c, hipMemsetAsync(c->wCtr.p, 0, 16 * sizeof(uint32_t), c->stream));
  HIPC(c, hipMemsetAsync(c->wCnt64.p, 0, 4 * sizeof(unsigned long long), c->stream));
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

// One batch of shots x frames through the bounce loop.
static int runBatch(ash_ctx* c, const BatchDesc& b, int maxDepth, const ash_render_opts* o, ash_stats* st) {
  SceneView sc = viewOf(c);
  hipStream_t s = c->stream;
  bool timing = o && o->timing;
  bool count = o && o->count_traversal;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (timing) {
    HIPC(c, hipEventCreate(&ev[0]));
    HIPC(c, hipEventCreate(&ev[1]));
  }
  auto timed = [&](double* acc, auto&& fn) -> int {
    if (!timing) {
      fn();
      return ASH_OK;
    }
    HIPC(c, hipEventRecord(ev[0], s));
    fn();
    HIPC(c, hipEventRecord(ev[1], s));
    HIPC(c, hipEventSynchronize(ev[1]));
    float ms = 0;
    HIPC(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
    *acc += ms;
    return ASH_OK;
  };
  launchRaygen(b, c->wf, s);
  uint32_t iters = o && o->max_bounce_iters ? o->max_bounce_iters : uint32_t(std::max(maxDepth, 1));
  int cur = 0;
  int rc = ASH_OK;
  for (uint32_t it = 0; it < iters && rc == ASH_OK; ++it) {
    launchResetCounters(c->wf, cur, s);
    rc = timed(&st->extend_ms, [&] { launchExtend(sc, c->wf, cur, b.n, c->traceGrid, count, s); });
    st->extend_launches++;
    if (rc) break;
    rc = timed(&st->shade_ms, [&] { launchShade(sc, b, c->wf, cur, b.n, s); });
    if (rc) break;
    rc = timed(&st->shadow_ms, [&] { launchShadow(sc, c->wf, b.n, c->traceGrid, count, s); });
    st->shadow_launches++;
    cur = 1 - cur;
  }
  if (rc == ASH_OK) launchAccumulate(b, c->wf, s);
  if (timing) {
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  HIPC(c, hipGetLastError());
  return rc;
}

static uint32_t framesPerBatch(uint32_t P, uint32_t shots, int32_t nFrames, const ash_render_opts* o) {
  uint64_t target = o && o->queue_paths ? o->queue_paths : (1u << 22);
  uint64_t perFrame = uint64_t(P) * shots;
  uint64_t f = std::max<uint64_t>(1, target / std::max<uint64_t>(perFrame, 1));
  return uint32_t(std::min<uint64_t>(f, uint64_t(nFrames)));
}

static int renderCommon(ash_ctx* c, uint32_t nShots, const ShotParams* hostShots, int32_t firstFrame, int32_t nFrames,
                        const ash_render_opts* o, ash_stats* st) {
  if (!c->hasScene || !c->hasAccel) return fail(c, ASH_E_STATE, "render called before upload_scene/build_accel");
  if (c->maxDepth > uint32_t(kStackLds + kSpill))
    return fail(c, ASH_E_UNSUPPORTED, "BVH depth %u exceeds the traversal stack (%d)", c->maxDepth, kStackLds + kSpill);
  if (nFrames <= 0 || firstFrame < 0) return fail(c, ASH_E_INVALID, "bad frame range");
  uint32_t P = hostShots[0].width * hostShots[0].height;
  int maxDepth = 0;
  for (uint32_t s = 0; s < nShots; ++s) {
    int rc = checkPc(c, &hostShots[s].pc);
    if (rc) return rc;
    maxDepth = std::max(maxDepth, hostShots[s].pc.maxPathDepth);
    if (hostShots[s].width * hostShots[s].height != P) return fail(c, ASH_E_INVALID, "shots must share one film size");
  }
  uint32_t F = framesPerBatch(P, nShots, nFrames, o);
  uint64_t cap64 = uint64_t(P) * nShots * F;
  if (cap64 > 0x7fffffffull) return fail(c, ASH_E_INVALID, "batch too large");
  int rc = ensureWavefront(c, uint32_t(cap64));
  if (rc) return rc;
  MALLOC(c, c->shotTab, nShots);
  HIPC(c, hipMemcpyAsync(c->shotTab.p, hostShots, nShots * sizeof(ShotParams), hipMemcpyHostToDevice, c->stream));
  ash_stats local{};
  ash_stats* S = st ? st : &local;
  std::memset(S, 0, sizeof *S);
  HIPC(c, hipMemsetAsync(c->wf.ctr + 5, 0, 4 * sizeof(uint32_t), c->stream));
  HIPC(c, hipMemsetAsync(c->wf.cnt64, 0, 4 * sizeof(unsigned long long), c->stream));
  hipEvent_t e0, e1;
  HIPC(c, hipEventCreate(&e0));
  HIPC(c, hipEventCreate(&e1));
  HIPC(c, hipEventRecord(e0, c->stream));
  uint64_t ext = 0, sh = 0, dsh = 0;
  for (int32_t f0 = firstFrame; f0 < firstFrame + nFrames; f0 += int32_t(F)) {
    BatchDesc b;
    b.P = P;
    b.width = hostShots[0].width;
    b.frames = uint32_t(std::min<int32_t>(int32_t(F), firstFrame + nFrames - f0));
    b.firstFrame = f0;
    b.shots = nShots;
    b.n = P * nShots * b.frames;
    b.shotTab = c->shotTab.p;
    rc = runBatch(c, b, maxDepth, o, S);
    if (rc) break;
    // ray counters are 32-bit on the device; fold them per batch
    uint32_t cnt[4];
    HIPC(c, hipMemcpyAsync(cnt, c->wf.ctr + 5, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipMemsetAsync(c->wf.ctr + 5, 0, 4 * sizeof(uint32_t), c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    ext += cnt[0];
    sh += cnt[1];
    dsh += cnt[2];
  }
  HIPC(c, hipEventRecord(e1, c->stream));
  HIPC(c, hipEventSynchronize(e1));
  float ms = 0;
  HIPC(c, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  S->render_ms = ms;
  S->paths = uint64_t(P) * nShots * uint64_t(nFrames);
  S->extension_rays = ext;
  S->shadow_rays = sh;
  S->dielectric_shadow_rays = dsh;
  if (o && o->count_traversal) {
    unsigned long long cnt[4];
    HIPC(c, hipMemcpy(cnt, c->wf.cnt64, sizeof cnt, hipMemcpyDeviceToHost));
    S->ext_node_visits = cnt[0];
    S->ext_tri_tests = cnt[1];
    S->shadow_node_visits = cnt[2];
    S->shadow_tri_tests = cnt[3];
  }
  return rc;
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

int ash_upload_scene(ash_ctx* c, const ash_scene_desc* d) {
  if (!c || !d) return ASH_E_INVALID;
  HIPC(c, hipSetDevice(c->device));
  c->hasScene = c->hasAccel = false;
  if (d->num_materials == 0 || d->num_lights == 0) return fail(c, ASH_E_INVALID, "material 0 and light 0 (dummies) are required");
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
      if (t != MaterialTypeBrdfLambertian && t != MaterialTypeBrdfPbrMetalnessRoughness &&
          t != MaterialTypeBrdfRoughPlastic && t != MaterialTypeBsdfDielectric && t != MaterialTypeBrdfEmissive)
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
  HIPC(c, hipMemcpyAsync(c->mats.p, d->materials, d->num_materials * sizeof(GpuMaterial), hipMemcpyHostToDevice, s));
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
    size_t npx = size_t(c->envW) * c->envH;
    MALLOC(c, c->env, npx);
    MALLOC(c, c->marg, c->envH);
    MALLOC(c, c->cond, npx);
    HIPC(c, hipMemcpyAsync(c->env.p, d->envmap.rgba, npx * sizeof(float4), hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->marg.p, marg.data(), marg.size() * 4, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->cond.p, cond.data(), cond.size() * 4, hipMemcpyHostToDevice, s));
  } else {
    c->envW = c->envH = 1;
    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 z2 = make_float2(0.f, 0.f);
    MALLOC(c, c->env, 1);
    MALLOC(c, c->marg, 1);
    MALLOC(c, c->cond, 1);
    HIPC(c, hipMemcpyAsync(c->env.p, &z, sizeof z, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->marg.p, &z2, sizeof z2, hipMemcpyHostToDevice, s));
    HIPC(c, hipMemcpyAsync(c->cond.p, &z2, sizeof z2, hipMemcpyHostToDevice, s));
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
  if (c->tris) (void)hipFree(c->tris), c->tris = nullptr;
  BuildArgs a;
  a.inst = c->inst.p;
  a.numInst = c->numInst;
  a.instTriOffset = c->instTriOffset.p;
  a.vertices = c->verts.p;
  a.indices = c->idx.p;
  a.numTris = c->numTris;
  a.leafMax = o && o->leaf_max_tris ? o->leaf_max_tris : 4;
  a.quality = o ? o->quality : 1;
  BuildResult r;
  char err[512] = {0};
  int rc = buildBvh(a, c->stream, r, err, sizeof err);
  if (rc) {
    if (r.nodes) (void)hipFree(r.nodes);
    if (r.tris) (void)hipFree(r.tris);
    return fail(c, rc, "%s", err);
  }
  c->nodes = r.nodes;
  c->tris = r.tris;
  c->nodeCount = r.numNodes;
  c->maxDepth = r.maxDepth;
  c->hasAccel = true;
  if (info) {
    info->num_triangles = c->numTris;
    info->num_nodes = r.numNodes;
    info->max_depth = r.maxDepth;
    info->build_ms = r.ms;
    info->sah_cost = r.sahCost;
    info->device_bytes = uint64_t(r.numNodes) * 64 + uint64_t(c->numTris) * 48 + c->verts.n * 16 + c->idx.n * 4;
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
  GpuPushConstantRaytrace pc;
  std::memset(&pc, 0, sizeof pc);
  pc.maxPathDepth = 1;
  ShotParams sp = makeShot(c, cam, &pc, nullptr, c->images.p, c->width, c->height);
  MALLOC(c, c->shotTab, 1);
  HIPC(c, hipMemcpyAsync(c->shotTab.p, &sp, sizeof sp, hipMemcpyHostToDevice, c->stream));
  DBuf<ash_hit> dout;
  MALLOC(c, dout, P);
  BatchDesc b;
  b.P = P;
  b.width = c->width;
  b.frames = 1;
  b.firstFrame = 0;
  b.shots = 1;
  b.n = P;
  b.shotTab = c->shotTab.p;
  SceneView sc = viewOf(c);
  launchRaygen(b, c->wf, c->stream);
  launchResetCounters(c->wf, 0, c->stream);
  launchExtend(sc, c->wf, 0, P, c->traceGrid, false, c->stream);
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
  HIPC(c, hipMemcpy(cd, c->cond.p, size_t(c->envW) * c->envH * sizeof(float2), hipMemcpyDeviceToHost));
  return ASH_OK;
}

void ash_pod_layout(uint32_t* o) {
  o[0] = sizeof(GpuVertex), o[1] = offsetof(GpuVertex, tangent);
  o[2] = sizeof(GpuMaterial), o[3] = offsetof(GpuMaterial, type);
  o[4] = sizeof(GpuLight), o[5] = offsetof(GpuLight, doubleSide);
  o[6] = sizeof(GpuCamera), o[7] = offsetof(GpuCamera, padding);
  o[8] = sizeof(GpuPushConstantRaytrace), o[9] = offsetof(GpuPushConstantRaytrace, uvOutChannel);
}

const char* ash_version(void) { return "asuna-mi355x 0.1 (gfx950, wavefront BVH2, LBVH+SAH collapse)"; }

}  // extern "C"
