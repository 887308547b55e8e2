// Device-resident scene layout (HBM):
//   instances  DevInstance[numInst]        128 B each (object<->world 3x4, mesh bases, ids)
//   vertices   float4[2*numVert]           (pos.xyz, uv.x), (normal.xyz, uv.y): 32 B/vertex
//   indices    uint32[numIdx]              per-mesh vertex indices, 3 per triangle
//   nodes      float4[4*numNodes]          BVH2, child boxes in the parent (64 B/node)
//   nodes4     float4[8*numNodes4]         4-wide BVH collapsed from it, one 128 B L2 line per
//                                           node: lo.x[4] hi.x[4] lo.y[4] hi.y[4] lo.z[4] hi.z[4]
//                                           child refs[4], unused[4]
//   tris       float4[3*numTris]           world-space v0,v1,v2 in leaf order (48 B/tri);
//                                           .w of v0/v1/v2 = instance, shade word, global id
//   materials  GpuMaterial[], lights GpuLight[], textures (RGBA32F), env + tables
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/asuna_hip.h"

namespace ashd {

struct DevInstance {
  float o2w[12];  // row-major 3x4 (gl_ObjectToWorldEXT)
  float w2o[12];  // row-major 3x4 (gl_WorldToObjectEXT)
  uint32_t vbase;     // first vertex of the mesh in `vertices`
  uint32_t ibase;     // first index of the mesh in `indices`
  uint32_t material;  // material id (unused for emitters)
  int32_t light;      // light id, -1 if not an emitter
  uint32_t triBase;   // global id of the instance's first triangle
  uint32_t ntri;
  uint32_t shade;     // shade word, carried by each triangle record into the hit buffer
  uint32_t pad;
};
// Shade word of an instance: bit 31 = surface of the second shading group,
// bit 30 = emitter, bits 24..28 = material type; bits 0..23 = light id
// (emitter) or material id.
constexpr uint32_t kShadeSide = 0x80000000u, kShadeLight = 0x40000000u, kShadeIdMask = 0x00ffffffu;
constexpr uint32_t kShadeTypeShift = 24, kShadeTypeMask = 0x1fu;

// Device copy of a material, padded to nine 16-byte words so that one can
// be fetched with nine aligned vector loads.
struct alignas(16) DevMaterial {
  GpuMaterial m;
  uint32_t pad[3];
};
static_assert(sizeof(DevMaterial) == 144, "DevMaterial layout");
static_assert(sizeof(DevInstance) == 128, "DevInstance layout");

// Surface plugins compiled into the second shading kernel (mirror,
// conductor, rough conductor, plastic, phong, kang18, disney).
__host__ __device__ inline bool isGroup2Type(uint32_t t) {
  return t == MaterialTypeBrdfMirror || t == MaterialTypeBrdfConductor || t == MaterialTypeBrdfRoughConductor ||
         t == MaterialTypeBrdfPlastic || t == MaterialTypeBrdfPhong || t == MaterialTypeBrdfKang18 ||
         t == MaterialTypeBrdfDisney;
}

// Leaf reference: ~((first << 3) | (count - 1)), count in [1, 8].
constexpr int kLeafMax = 8;
__host__ __device__ inline int32_t encodeLeaf(uint32_t first, uint32_t count) {
  return ~int32_t((first << 3) | (count - 1));
}
__host__ __device__ inline void decodeLeaf(int32_t ref, uint32_t& first, uint32_t& count) {
  uint32_t v = uint32_t(~ref);
  first = v >> 3;
  count = (v & 7u) + 1u;
}
// A child slot that holds nothing: far-away degenerate box and this ref.
constexpr int32_t kEmptyRef = int32_t(0x7fffffff);

// Two-level acceleration (instancing without duplication): one object-space
// BVH (BLAS) per distinct mesh, shared by all its instances, under a binary
// tree over the instances' world boxes (TLAS).  A BLAS triangle record keeps
// the mesh-local primitive in v2.w; the instance, its shade word and its
// global triangle ids come from the instance walked.
struct DevBlas {
  const float4* nodes4;  // 4-wide nodes, object space
  const float4* tris;    // 3 float4 per triangle in the BLAS's leaf order
  uint32_t shadeBase;    // first shading record of the BLAS in the shared array
  uint32_t ntri;
};

struct BuildArgs {
  const DevInstance* inst;
  uint32_t numInst;
  const uint32_t* instTriOffset;  // numInst + 1 prefix of triangle counts
  const float4* vertices;         // 2 float4 per vertex
  const uint32_t* indices;
  uint32_t numTris;
  uint32_t leafMax;   // 1..8
  uint32_t quality;   // 0 = LBVH (one triangle per leaf), 1 = SAH leaf collapse
  uint32_t branching; // 2 or 4
  uint32_t quantize;  // 1: also emit quantized 64-byte 4-wide nodes (branching 4)
};

struct BuildResult {
  float4* nodes = nullptr;   // device, 4 float4 per binary node
  float4* nodes4 = nullptr;  // device, 8 float4 per 4-wide node (branching 4)
  float4* nodesQ = nullptr;  // device, 4 float4 per quantized 4-wide node (same ids and refs as nodes4)
  float4* tris = nullptr;    // device, 3 float4 per triangle
  float4* shade = nullptr;   // device, 6 float4 per triangle (leaf order): object-space vertex pairs
  uint32_t numNodes = 0;
  uint32_t maxDepth = 0;
  uint32_t numNodes4 = 0;
  uint32_t depth4 = 0;
  double sahCost = 0;
  double ms = 0;
};

// Builds the BVH on `stream`; allocations in result are owned by the caller.
// Returns 0 or a negative ASH_E_* code with a message in err.
int buildBvh(const BuildArgs& a, hipStream_t stream, BuildResult& r, char* err, size_t errLen);

}  // namespace ashd
