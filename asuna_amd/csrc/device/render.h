// Wavefront render path: data shared by render.hip (kernels) and
// context.hip (host orchestration).
#pragma once

#include "devcommon.h"
#include "devscene.h"

namespace ashd {

// Per-shot constants (camera UBO + push constants + output planes).
struct ShotParams {
  GpuCamera cam;
  GpuPushConstantRaytrace pc;
  int32_t depthChannel;  // extension AOV channel receiving first-hit t
  uint32_t width, height;
  uint32_t pad;
  float4* images;  // 9 planes of width*height float4
};

struct SceneView {
  const float4* nodes;
  const float4* tris;
  const DevInstance* inst;
  const float4* verts;
  const uint32_t* idx;
  const GpuMaterial* mats;
  const GpuLight* lights;
  const DevImage* textures;
  uint32_t numTextures;
  uint32_t hasEnv;
  DevImage env;
  const float2* marg;  // H entries (column 0 of the marginal table)
  const float2* cond;  // W*H entries
};

// Wavefront state for up to `capacity` paths.  Path p of a batch covers
// shot = p / (frames*P), frame slot = (p / P) % frames, pixel = p % P.
struct Wavefront {
  uint32_t capacity;
  float4* ro;    // ray origin of the current segment
  float4* rd;    // ray direction
  float4* thr;   // throughput.xyz, w = pdf of the last BSDF sample
  float4* rad;   // radiance.xyz, w = Gaussian filter weight
  float4* ld;    // pending NEE radiance of the current bounce
  uint4* info;   // x seed, y depth | bsdfFlags << 16, z pixel, w frame | shot << 16
  float4* hitF;  // t, u, v
  int4* hitI;    // instance (-1 miss), primitive, global triangle id, leaf position
  int* queue[2];
  float4* shO;  // shadow rays: origin.xyz, tmax
  float4* shD;  // direction.xyz, path id
  uint32_t* ctr;  // [0],[1] queue counts, [2] shadow count, [3] extend fetch,
                  // [4] shadow fetch, [5..8] stats: ext rays, shadow rays,
                  // dielectric shadow rays, paths
  int* spill;     // traversal stack spill, kSpill ints per persistent thread
  uint32_t spillThreads;
  unsigned long long* cnt64;  // [0] ext nodes, [1] ext tris, [2] shadow nodes, [3] shadow tris
};

struct BatchDesc {
  uint32_t P;          // pixels per shot
  uint32_t width;
  uint32_t frames;     // frames per shot in this batch
  int32_t firstFrame;  // frame index of slot 0
  uint32_t shots;      // shots in this batch
  uint32_t n;          // = P * frames * shots
  const ShotParams* shotTab;  // device array indexed by shot slot
};

constexpr int kStackLds = 16;    // per-thread traversal stack entries in LDS
constexpr int kSpill = 48;       // further entries spilled to global memory
constexpr int kTraceBlock = 256;

// Host-side launchers (render.hip).
void launchRaygen(const BatchDesc& b, const Wavefront& w, hipStream_t s);
void launchExtend(const SceneView& sc, const Wavefront& w, int cur, uint32_t maxRays, uint32_t grid,
                  bool count, hipStream_t s);
void launchShade(const SceneView& sc, const BatchDesc& b, const Wavefront& w, int cur, uint32_t maxRays,
                 hipStream_t s);
void launchShadow(const SceneView& sc, const Wavefront& w, uint32_t maxRays, uint32_t grid, bool count,
                  hipStream_t s);
void launchAccumulate(const BatchDesc& b, const Wavefront& w, hipStream_t s);
void launchResetCounters(const Wavefront& w, int cur, hipStream_t s);
void launchFirstHit(const SceneView& sc, const ShotParams* shot, const Wavefront& w, ash_hit* out, uint32_t P,
                    hipStream_t s);

}  // namespace ashd
