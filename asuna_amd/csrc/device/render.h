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

// Sun & sky record plus its direction-independent terms (sun direction
// after the horizon clamp, unit scale, sun colours, ground colour), which
// kSkyPrepare computes once per record on the device.
struct DevSky {
  GpuSunAndSky ss;
  float sun[3];      // tweaked sun direction, clamped to z >= 0.001
  float realSun[3];  // tweaked sun direction before the clamp
  float rgbScale[3];
  float sunUp[3];    // calc_sun_color at the local haze (view above the horizon)
  float sunDown[3];  // calc_sun_color at haze 2 (view at or below the horizon)
  float ground[3];   // ground colour term of the below-horizon branch
  float factor;      // night factor of the sky tint (1 by day)
  float haze, sat, hh;
  float diskScale, glowScale;  // calc_physical_scale (1, 1 when not physically scaled)
  int32_t inUse;
  int32_t pad;
};

struct SceneView {
  const float4* nodes;   // BVH2 (always built)
  const float4* nodes4;  // 4-wide BVH, or null to traverse the BVH2
  const float4* nodesQ;  // quantized 64-byte 4-wide nodes, or null (same refs as nodes4)
  const float4* tris;
  const float4* shadeTris;  // 6 float4 per triangle in leaf order (object-space vertex pairs)
  const DevInstance* inst;
  const float4* verts;
  const uint32_t* idx;
  const DevMaterial* mats;
  uint32_t numMats;
  const GpuLight* lights;
  const DevImage* textures;
  uint32_t numTextures;
  uint32_t hasEnv;
  DevImage env;        // W*H texels: (r, g, b, conditional pdf) in 4x2-texel tiles; the image's alpha is unused
  const float2* marg;  // H entries (column 0 of the marginal table)
  const float* condX;  // W*H entries: the conditional table's inverse-CDF channel, in 8x4-texel tiles
  const DevSky* sky;   // sun & sky (inUse 0: env map / background)
  // two-level acceleration (twoLevel = 1): TLAS binary nodes over instances
  // (leaf = one instance: its index in tlasInst), per-instance BLAS slot
  uint32_t twoLevel;
  const float4* tlas;
  const uint32_t* tlasInst;
  const DevBlas* blas;
  const uint32_t* instBlas;
};

// Wavefront state for up to `capacity` paths (path numbering: WorkUnit).
// A path's radiance is indexed by path id (the accumulation reads it
// there); its ray, throughput and RNG / pixel record, and the hit of its
// ray, live in the slot order of the extension queue that holds it, so that
// traversal and shading read and write them in runs (the queue after
// raygen is the identity: slot = path id).
struct Wavefront {
  uint32_t capacity;
  float4* rayO[2];  // extension queues (alternating): ray origin.xyz per slot
  float4* rayD[2];  // ray direction.xyz, w = path id
  float4* thr[2];   // throughput.xyz, w = pdf of the last BSDF sample (read only at depth > 1)
  uint4* info[2];   // x seed, y depth | bsdfFlags << 16, z pixel, w shot | (frame == 0) << 16
  float4* rad;   // by path id: radiance.xyz, w = Gaussian filter weight
  float4* ld;    // pending NEE radiance of the current bounce, per shadow-ray slot
  float4* hitF;  // per slot of the queue being traced: t, u, v
  int4* hitI;    // instance (-1 miss), shade word, global triangle id, leaf position
  int* side;     // second-group shading queue: slots of the current extension queue
  float4* shO;  // shadow rays: origin.xyz, tmax
  float4* shD;  // direction.xyz, path id
  uint32_t* ctr;  // [0],[1] queue counts, [2] shadow count, [3] extend fetch,
                  // [4] shadow fetch, [8] second-group shading queue count,
                  // [kXcdCtr + 32 j]: fetch counters of the per-XCD queue slices
                  // (j < 8 extend, j >= 8 shadow), one 128-byte line each
  int* spill;       // traversal stack spill: spillCap entries per persistent thread
  float* spillT;    // entry distances of the spilled entries (4-wide closest hit)
  uint32_t spillThreads;
  uint32_t spillCap;
  unsigned long long* cnt64;  // [0] ext nodes, [1] ext tris, [2] shadow nodes, [3] shadow tris,
                              // [4] extension rays, [5] shadow rays, [6] dielectric shadow rays,
                              // [7..12] of the fused launches: ext nodes, ext tris, shadow nodes,
                              // shadow tris, ext rays, shadow rays (kNumCnt64 in all)
};

// A work unit: frames [f0, f0 + nf) of the pixel run [pix0, pix0 + np) of
// one shot (positions in the film's tile order).  Its paths are
// base + k * nf + fi (pixel-major: a wave's paths are a few pixels' frames,
// whose camera rays nearly coincide), and its accumulate threads accBase + k.  The units of one batch cover disjoint
// (shot, pixel) sets, so each pixel's frames are folded by one thread in
// frame order.
struct WorkUnit {
  uint32_t shot;  // slot in the shot table
  uint32_t pix0, np;
  int32_t f0;
  uint32_t nf;
  uint32_t base;
  uint32_t accBase;
  uint32_t pad;
};

struct BatchDesc {
  uint32_t n;       // paths (= sum of np * nf)
  uint32_t nAcc;    // accumulate threads (= sum of np)
  uint32_t nUnits;  // >= 1
  uint32_t shots;   // 1: every path belongs to the shot copied into `one` (kernel-argument
                    // constants); any other value: shading reads the shot table
  const WorkUnit* units;      // device array of nUnits
  const ShotParams* shotTab;  // device array indexed by shot slot
  ShotParams one;             // kernel-argument copy of the only shot (scalar loads)
};

constexpr int kStackLds = 12;    // per-thread traversal stack entries in LDS (default)
constexpr int kStackLdsMin = 8;  // smallest LDS share (the spill covers the rest)
constexpr int kSpill = 48;       // minimum further entries spilled to global memory
constexpr int kTraceBlock = 256;
constexpr int kXcds = 8;          // XCDs of the device; workgroup b runs on XCD b % 8
constexpr int kXcdCtr = 32;       // first per-slice fetch counter in Wavefront::ctr
constexpr int kNumCtr = kXcdCtr + 2 * kXcds * 32;
constexpr int kNumCnt64 = 16;

// Ray fetch scheme of the persistent traversal kernels.
enum TraceScheme : uint32_t {
  kSchemeXcdSlices = 0,  // wave fetch from the wave's own XCD slice of the queue first
  kSchemeQueueOrder = 1, // wave fetch in queue order
  kSchemeRefill = 2,     // per-lane refill (4-wide tree only)
};

// Host-side launchers (render.hip).
void launchRaygen(const BatchDesc& b, const Wavefront& w, hipStream_t s);
// ldsStack: traversal stack entries per thread kept in LDS (8 or 12) by the
// 4-wide closest-hit kernel
void launchExtend(const SceneView& sc, const Wavefront& w, int cur, uint32_t ldsStack, uint32_t scheme,
                  uint32_t grid, bool count, hipStream_t s);
// Closest hit of the camera rays (queue 0 after raygen): the wave-coherent
// walk on the float 4-wide tree with XCD slices, else launchExtend.
// coarse: a wave fetches 256 rays per atomic instead of 64 (for trees of at
// most kCoarseFetchTris triangles, whose cheap rays would otherwise wait on
// the fetch atomics).
constexpr uint32_t kCoarseFetchTris = 1u << 16;
bool canPacketCamera(const SceneView& sc, uint32_t scheme);
void launchExtendCamera(const SceneView& sc, const Wavefront& w, uint32_t ldsStack, uint32_t scheme, uint32_t grid,
                        bool count, bool coarse, hipStream_t s);
void launchShade(const SceneView& sc, const BatchDesc& b, const Wavefront& w, int cur, uint32_t maxRays,
                 hipStream_t s);
void launchShadow(const SceneView& sc, const Wavefront& w, uint32_t scheme, uint32_t grid, bool count,
                  hipStream_t s);
void launchAccumulate(const BatchDesc& b, const Wavefront& w, hipStream_t s);
void launchResetCounters(const Wavefront& w, int cur, hipStream_t s);
// Resets of the fused loop: the counts shading appends to (and the shadow
// fetch counters), and the extension fetch counters.
void launchResetShade(const Wavefront& w, int cur, hipStream_t s);
void launchResetExtend(const Wavefront& w, hipStream_t s);
// Shadow rays of the current bounce and extension rays of queue `cur` in
// one persistent launch (default traversal scheme on the float 4-wide tree).
bool canFuseTrace(const SceneView& sc, uint32_t scheme);
void launchTrace(const SceneView& sc, const Wavefront& w, int cur, uint32_t ldsStack, uint32_t grid, bool count,
                 hipStream_t s);
// Fills the derived terms of *sky from sky->ss (one thread).
void launchSkyPrepare(DevSky* sky, hipStream_t s);
void launchFirstHit(const SceneView& sc, const ShotParams* shot, const Wavefront& w, ash_hit* out, uint32_t P,
                    hipStream_t s);

}  // namespace ashd
