// GPU BVH construction for gfx950 (replaces the driver BLAS/TLAS build,
// pipeline_raytrace.cpp:107-147).
//
// One flat BVH over world-space triangles of all instances:
//   1. world triangles + boxes + centroids (one thread per triangle)
//   2. centroid bounds (block reduction)
//   3. 63-bit Morton keys (21 bits/axis), stable LSD radix sort (below)
//   4. binary radix tree, Karras 2012 "Maximizing Parallelism in the
//      Construction of BVHs, Octrees, and k-d Trees" (equal keys ordered by
//      index)
//   5. bottom-up refit with arrival counters: boxes, triangle counts and the
//      SAH cost; a subtree with <= leafMax triangles whose leaf cost does not
//      exceed its split cost is collapsed into one leaf
//   6. compaction of the surviving internal nodes (prefix sum) and emission
//      of 64-byte BVH2 nodes that hold both child boxes; triangles are
//      reordered into leaf order.
#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <vector>

#include "devscene.h"

namespace ashd {

namespace {

constexpr int kBlock = 256;
constexpr float kCostNode = 1.0f;
constexpr float kCostTri = 1.0f;

#define HIPB(x)                                                                    \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::snprintf(err, errLen, "bvh: %s failed: %s", #x, hipGetErrorString(e_)); \
      return ASH_E_HIP;                                                            \
    }                                                                              \
  } while (0)

__device__ __forceinline__ float3 xform(const float* m, float4 p) {
  // ((m0*x + m1*y) + m2*z) + m3 per row
  return make_float3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                     m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}

__global__ void kWorldTris(BuildArgs a, float4* wtri, float4* boxLo, float4* boxHi) {
  uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.numTris) return;
  // instance by binary search over the triangle-count prefix
  uint32_t lo = 0, hi = a.numInst;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (a.instTriOffset[mid] <= g) lo = mid;
    else hi = mid;
  }
  const DevInstance& in = a.inst[lo];
  uint32_t prim = g - a.instTriOffset[lo];
  float3 p[3];
  for (int k = 0; k < 3; ++k) {
    uint32_t vi = in.vbase + a.indices[in.ibase + prim * 3 + k];
    p[k] = xform(in.o2w, a.vertices[2 * size_t(vi)]);
  }
  wtri[3 * size_t(g) + 0] = make_float4(p[0].x, p[0].y, p[0].z, __int_as_float(int(lo)));
  wtri[3 * size_t(g) + 1] = make_float4(p[1].x, p[1].y, p[1].z, __uint_as_float(in.shade));
  wtri[3 * size_t(g) + 2] = make_float4(p[2].x, p[2].y, p[2].z, __int_as_float(int(g)));
  boxLo[g] = make_float4(fminf(p[0].x, fminf(p[1].x, p[2].x)), fminf(p[0].y, fminf(p[1].y, p[2].y)),
                         fminf(p[0].z, fminf(p[1].z, p[2].z)), 0.f);
  boxHi[g] = make_float4(fmaxf(p[0].x, fmaxf(p[1].x, p[2].x)), fmaxf(p[0].y, fmaxf(p[1].y, p[2].y)),
                         fmaxf(p[0].z, fmaxf(p[1].z, p[2].z)), 0.f);
}

__global__ void kCentroidBounds(const float4* boxLo, const float4* boxHi, uint32_t n, float4* partLo, float4* partHi) {
  __shared__ float sl[3][kBlock], sh[3][kBlock];
  float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 a = boxLo[i], b = boxHi[i];
    float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
    for (int k = 0; k < 3; ++k) l[k] = fminf(l[k], c[k]), h[k] = fmaxf(h[k], c[k]);
  }
  for (int k = 0; k < 3; ++k) sl[k][threadIdx.x] = l[k], sh[k][threadIdx.x] = h[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s)
      for (int k = 0; k < 3; ++k) {
        sl[k][threadIdx.x] = fminf(sl[k][threadIdx.x], sl[k][threadIdx.x + s]);
        sh[k][threadIdx.x] = fmaxf(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partLo[blockIdx.x] = make_float4(sl[0][0], sl[1][0], sl[2][0], 0.f);
    partHi[blockIdx.x] = make_float4(sh[0][0], sh[1][0], sh[2][0], 0.f);
  }
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

__global__ void kMorton(const float4* boxLo, const float4* boxHi, uint32_t n, const float4* partLo,
                        const float4* partHi, uint32_t nPart, uint64_t* keys, uint32_t* ids) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (uint32_t p = 0; p < nPart; ++p) {
    float4 a = partLo[p], b = partHi[p];
    l[0] = fminf(l[0], a.x), l[1] = fminf(l[1], a.y), l[2] = fminf(l[2], a.z);
    h[0] = fmaxf(h[0], b.x), h[1] = fmaxf(h[1], b.y), h[2] = fmaxf(h[2], b.z);
  }
  float4 a = boxLo[i], b = boxHi[i];
  float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
  uint64_t key = 0;
  for (int k = 0; k < 3; ++k) {
    float ext = h[k] - l[k];
    float t = ext > 0.f ? (c[k] - l[k]) / ext : 0.5f;
    t = fminf(fmaxf(t, 0.f), 1.f);
    uint32_t q = uint32_t(fminf(t * 2097152.0f, 2097151.0f));
    key |= spread21(q) << (2 - k);
  }
  keys[i] = key;
  ids[i] = i;
}

__device__ __forceinline__ int deltaK(const uint64_t* k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint64_t a = k[i], b = k[j];
  if (a == b) return 64 + __clz(uint32_t(i ^ j));
  return __clzll(a ^ b);
}

// Internal node i in [0, n-2]; children: c >= 0 internal, c < 0 leaf ~c.
__global__ void kKarras(const uint64_t* keys, int n, int* left, int* right, int* parent, int* leafParent,
                        int2* range) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (deltaK(keys, n, i, i + 1) - deltaK(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = deltaK(keys, n, i, i - d);
  int lmax = 2;
  while (deltaK(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (deltaK(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = deltaK(keys, n, i, j);
  int s = 0, step = l;
  do {
    step = (step + 1) >> 1;
    int ns = s + step;
    if (ns < l && deltaK(keys, n, i, i + ns * d) > dnode) s = ns;
  } while (step > 1);
  int gamma = i + s * d + min(d, 0);
  int lo = min(i, j), hi = max(i, j);
  int lc = (lo == gamma) ? ~gamma : gamma;
  int rc = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
  left[i] = lc;
  right[i] = rc;
  range[i] = make_int2(lo, hi);
  if (lc >= 0) parent[lc] = i;
  else leafParent[~lc] = i;
  if (rc >= 0) parent[rc] = i;
  else leafParent[~rc] = i;
  if (i == 0) parent[0] = -1;
}

__device__ __forceinline__ float halfArea(float4 lo, float4 hi) {
  float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  return dx * dy + dy * dz + dz * dx;
}

// ----------------------------------------------------------------- PLOC
// Parallel locally-ordered clustering (Meister and Bittner 2018): clusters
// kept in Morton order; each finds, among the kPlocRadius clusters on either
// side, the one whose merged box has the smallest surface area (ties to the
// smaller index, so the globally best pair is always mutual); mutual pairs
// merge into a new internal node; the survivors are compacted in order.
// Internal ids are handed out from n - 2 down, so the root is node 0.
constexpr int kPlocRadius = 16;

__global__ void kPlocInit(uint32_t n, const uint32_t* sortedIds, const float4* boxLo, const float4* boxHi, int* cNode,
                          float4* cLo, float4* cHi) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t t = sortedIds[i];
  cNode[i] = ~int(i);
  cLo[i] = boxLo[t];
  cHi[i] = boxHi[t];
}

__global__ __launch_bounds__(kBlock) void kPlocNearest(uint32_t m, const float4* cLo, const float4* cHi, int* nn) {
  __shared__ float4 sl[kBlock + 2 * kPlocRadius], sh[kBlock + 2 * kPlocRadius];
  const int base = int(blockIdx.x * blockDim.x) - kPlocRadius;
  for (int k = int(threadIdx.x); k < kBlock + 2 * kPlocRadius; k += kBlock) {
    int g = base + k;
    if (g >= 0 && g < int(m)) {
      sl[k] = cLo[g];
      sh[k] = cHi[g];
    }
  }
  __syncthreads();
  const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= int(m)) return;
  const int li = i - base;
  const float4 a = sl[li], b = sh[li];
  float best = FLT_MAX;
  int bj = -1;
  const int j0 = max(0, i - kPlocRadius), j1 = min(int(m) - 1, i + kPlocRadius);
  for (int j = j0; j <= j1; ++j) {
    if (j == i) continue;
    const float4 c = sl[j - base], d = sh[j - base];
    float4 lo = make_float4(fminf(a.x, c.x), fminf(a.y, c.y), fminf(a.z, c.z), 0.f);
    float4 hi = make_float4(fmaxf(b.x, d.x), fmaxf(b.y, d.y), fmaxf(b.z, d.z), 0.f);
    float cost = halfArea(lo, hi);
    if (cost < best) best = cost, bj = j;  // ascending j: ties keep the smaller index
  }
  nn[i] = bj;
}

__global__ void kPlocMerge(uint32_t m, uint32_t n, const int* nn, const int* cNode, const float4* cLo, const float4* cHi,
                           int* oNode, float4* oLo, float4* oHi, uint32_t* valid, int* left, int* right, int* parent,
                           int* leafParent, uint32_t* nodeCnt, unsigned* idCounter) {
  int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= int(m)) return;
  int j = nn[i];
  bool mutual = j >= 0 && nn[j] == i;
  if (mutual && i > j) {  // absorbed by its partner
    valid[i] = 0;
    return;
  }
  valid[i] = 1;
  if (!mutual) {
    oNode[i] = cNode[i];
    oLo[i] = cLo[i];
    oHi[i] = cHi[i];
    return;
  }
  int id = int(n - 2) - int(atomicAdd(idCounter, 1u));
  int c0 = cNode[i], c1 = cNode[j];
  left[id] = c0;
  right[id] = c1;
  uint32_t cnt = 0;
  for (int c : {c0, c1}) {
    if (c >= 0) {
      parent[c] = id;
      cnt += nodeCnt[c];
    } else {
      leafParent[~c] = id;
      cnt += 1u;
    }
  }
  nodeCnt[id] = cnt;
  float4 a = cLo[i], b = cHi[i], c = cLo[j], d = cHi[j];
  oNode[i] = id;
  oLo[i] = make_float4(fminf(a.x, c.x), fminf(a.y, c.y), fminf(a.z, c.z), 0.f);
  oHi[i] = make_float4(fmaxf(b.x, d.x), fmaxf(b.y, d.y), fmaxf(b.z, d.z), 0.f);
}

__global__ void kPlocCompact(uint32_t m, const uint32_t* valid, const uint32_t* offs, const int* oNode,
                             const float4* oLo, const float4* oHi, int* cNode, float4* cLo, float4* cHi) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !valid[i]) return;
  uint32_t o = offs[i];
  cNode[o] = oNode[i];
  cLo[o] = oLo[i];
  cHi[o] = oHi[i];
}

// Leaves in depth-first order, so each subtree covers one run of positions:
// the new position of a leaf is the number of leaves left of it, summed
// over the ancestors whose right subtree holds it.
__global__ void kPlocLeafOrder(uint32_t n, const int* left, const int* right, const int* parent, const int* leafParent,
                               const uint32_t* nodeCnt, const uint32_t* sortedIds, uint32_t* posOf,
                               uint32_t* newSorted, int* newLeafParent) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int child = ~int(p);
  int node = leafParent[p];
  uint32_t pos = 0;
  while (node >= 0) {
    if (right[node] == child) {
      int l = left[node];
      pos += l >= 0 ? nodeCnt[l] : 1u;
    }
    child = node;
    node = parent[node];
  }
  posOf[p] = pos;
  newSorted[pos] = sortedIds[p];
  newLeafParent[pos] = leafParent[p];
}

// Leaf range of each internal node (its leftmost leaf's new position and
// its leaf count); runs before the child refs are relabelled.
__global__ void kPlocRanges(uint32_t nInternal, const int* left, const uint32_t* posOf, const uint32_t* nodeCnt,
                            int2* range) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nInternal) return;
  int k = int(i);
  while (left[k] >= 0) k = left[k];
  uint32_t first = posOf[~left[k]];
  range[i] = make_int2(int(first), int(first + nodeCnt[i] - 1));
}

__global__ void kPlocRelabel(uint32_t nInternal, int* left, int* right, const uint32_t* posOf) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nInternal) return;
  int l = left[i], r = right[i];
  if (l < 0) left[i] = ~int(posOf[~l]);
  if (r < 0) right[i] = ~int(posOf[~r]);
}

// Bottom-up boxes, SAH costs and leaf-collapse decisions, one tree level
// per launch (deepest first): the launch boundary orders a node's reads
// after its children's writes, so no device-scope fences are needed (on
// this chip each one writes back or invalidates the XCD's L2).
// level[i] = depth of internal node i (root 0); maxLevel = deepest level
// (nInternal or more only if the parent links hold a cycle).
__global__ void kNodeLevel(int nInternal, const int* parent, int* level, unsigned* maxLevel) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int d = 0;
  if (i < nInternal) {
    for (int p = parent[i]; p >= 0 && d < nInternal; p = parent[p]) ++d;  // bounded: a cycle ends the walk
    level[i] = d;
  }
  for (int off = 32; off > 0; off >>= 1) d = max(d, __shfl_xor(d, off));
  if (__lane_id() == 0 && d > 0) atomicMax(maxLevel, unsigned(d));
}

__global__ void kRefitLevel(int nInternal, int lv, const int* level, const uint32_t* sortedIds, const float4* boxLo,
                            const float4* boxHi, const int* left, const int* right, const int2* range, float4* nodeLo,
                            float4* nodeHi, float* cost, uint8_t* collapsed, uint32_t leafMax, uint32_t quality) {
  int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nInternal || level[node] != lv) return;
  int c[2] = {left[node], right[node]};
  float4 lo = make_float4(FLT_MAX, FLT_MAX, FLT_MAX, 0.f), hi = make_float4(-FLT_MAX, -FLT_MAX, -FLT_MAX, 0.f);
  float childCost = 0.f;
  for (int q = 0; q < 2; ++q) {
    float4 a, b;
    float cc;
    if (c[q] < 0) {
      uint32_t t = sortedIds[~c[q]];
      a = boxLo[t];
      b = boxHi[t];
      cc = kCostTri * halfArea(a, b);
    } else {
      a = nodeLo[c[q]];
      b = nodeHi[c[q]];
      cc = cost[c[q]];
    }
    lo = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), 0.f);
    hi = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), 0.f);
    childCost += cc;
  }
  float area = halfArea(lo, hi);
  int2 rg = range[node];
  uint32_t cnt = uint32_t(rg.y - rg.x + 1);
  float splitCost = kCostNode * area + childCost;
  float leafCost = kCostTri * area * float(cnt);
  bool col = quality > 0 && cnt <= leafMax && leafCost <= splitCost;
  nodeLo[node] = lo;
  nodeHi[node] = hi;
  cost[node] = col ? leafCost : splitCost;
  collapsed[node] = col ? 1 : 0;
}

// emitted[i] = internal node i survives (neither it nor an ancestor collapsed).
__global__ void kEmitFlags(int nInternal, const int* parent, const uint8_t* collapsed, uint32_t* emitted,
                           unsigned* maxDepth) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned deepest = 0;  // one atomic per wave
  if (i < nInternal) {
    bool e = !collapsed[i];
    unsigned depth = 1;
    for (int p = parent[i]; p >= 0 && e; p = parent[p]) {
      if (collapsed[p]) e = false;
      ++depth;
    }
    emitted[i] = e ? 1u : 0u;
    if (e) deepest = depth;
  }
  for (int off = 32; off > 0; off >>= 1) deepest = max(deepest, unsigned(__shfl_xor(int(deepest), off)));
  if (__lane_id() == 0 && deepest > 0) atomicMax(maxDepth, deepest);
}

__device__ __forceinline__ int32_t childRef(int c, const int2* range, const uint8_t* collapsed, const uint32_t* newIdx) {
  if (c < 0) return encodeLeaf(uint32_t(~c), 1u);
  if (collapsed[c]) {
    int2 rg = range[c];
    return encodeLeaf(uint32_t(rg.x), uint32_t(rg.y - rg.x + 1));
  }
  return int32_t(newIdx[c]);
}

__global__ void kEmitNodes(int nInternal, const uint32_t* emitted, const uint32_t* newIdx, const int* left,
                           const int* right, const int2* range, const uint8_t* collapsed, const uint32_t* sortedIds,
                           const float4* boxLo, const float4* boxHi, const float4* nodeLo, const float4* nodeHi,
                           float4* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nInternal || !emitted[i]) return;
  int c[2] = {left[i], right[i]};
  float4 lo[2], hi[2];
  for (int q = 0; q < 2; ++q) {
    if (c[q] < 0) {
      uint32_t t = sortedIds[~c[q]];
      lo[q] = boxLo[t];
      hi[q] = boxHi[t];
    } else {
      lo[q] = nodeLo[c[q]];
      hi[q] = nodeHi[c[q]];
    }
  }
  size_t o = size_t(newIdx[i]) * 4;
  out[o + 0] = make_float4(lo[0].x, hi[0].x, lo[0].y, hi[0].y);
  out[o + 1] = make_float4(lo[1].x, hi[1].x, lo[1].y, hi[1].y);
  out[o + 2] = make_float4(lo[0].z, hi[0].z, lo[1].z, hi[1].z);
  out[o + 3] = make_float4(__int_as_float(childRef(c[0], range, collapsed, newIdx)),
                           __int_as_float(childRef(c[1], range, collapsed, newIdx)), 0.f, 0.f);
}

__global__ void kReorderTris(uint32_t n, const uint32_t* sortedIds, const float4* wtri, float4* tris) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t t = sortedIds[j];
  for (int k = 0; k < 3; ++k) tris[3 * size_t(j) + k] = wtri[3 * size_t(t) + k];
}

// Shading records in leaf order: the object-space vertex pairs (position +
// u, normal + v) of the three corners of each triangle, so a hit reads one
// contiguous 96-byte record instead of an index triple and three vertices.
__global__ void kShadeTris(uint32_t n, const uint32_t* sortedIds, BuildArgs a, const float4* wtri, float4* shade) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t t = sortedIds[j];
  uint32_t inst = uint32_t(__float_as_int(wtri[3 * size_t(t)].w));
  const DevInstance& in = a.inst[inst];
  uint32_t prim = t - in.triBase;  // t is the global triangle id
  for (int k = 0; k < 3; ++k) {
    uint32_t vi = in.vbase + a.indices[in.ibase + prim * 3 + k];
    shade[6 * size_t(j) + 2 * k] = a.vertices[2 * size_t(vi)];
    shade[6 * size_t(j) + 2 * k + 1] = a.vertices[2 * size_t(vi) + 1];
  }
}

// Root for scenes the radix tree cannot express (0 or 1 triangle, or the
// whole tree collapsed into one leaf): child 0 = one leaf, child 1 = empty.
__global__ void kSingleLeafRoot(uint32_t n, const float4* boxLo, const float4* boxHi, const uint32_t* sortedIds,
                                float4* out) {
  float4 lo = make_float4(FLT_MAX, FLT_MAX, FLT_MAX, 0.f), hi = make_float4(-FLT_MAX, -FLT_MAX, -FLT_MAX, 0.f);
  for (uint32_t j = 0; j < n; ++j) {
    uint32_t t = sortedIds[j];
    float4 a = boxLo[t], b = boxHi[t];
    lo = make_float4(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z), 0.f);
    hi = make_float4(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z), 0.f);
  }
  const float far = 3.0e38f;
  int32_t r0 = n ? encodeLeaf(0, n) : kEmptyRef;
  if (!n) lo = hi = make_float4(far, far, far, 0.f);
  out[0] = make_float4(lo.x, hi.x, lo.y, hi.y);
  out[1] = make_float4(far, far, far, far);
  out[2] = make_float4(lo.z, hi.z, far, far);
  out[3] = make_float4(__int_as_float(r0), __int_as_float(kEmptyRef), 0.f, 0.f);
}

// Structural check of the emitted BVH before any traversal touches it:
// every reference is in range, leaves address valid triangles, and every
// node except the root has exactly one parent (so traversal from the root
// visits a tree and terminates).  bad[0] counts violations.
__global__ void kValidate(const float4* nodes, uint32_t numNodes, uint32_t numTris, unsigned* refcnt, unsigned* bad) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numNodes) return;
  float4 n3 = nodes[size_t(i) * 4 + 3];
  int32_t refs[2] = {__float_as_int(n3.x), __float_as_int(n3.y)};
  for (int q = 0; q < 2; ++q) {
    int32_t r = refs[q];
    if (r == kEmptyRef) continue;
    if (r >= 0) {
      if (uint32_t(r) >= numNodes || r == 0) atomicAdd(bad, 1u);
      else atomicAdd(&refcnt[r], 1u);
    } else {
      uint32_t first, cnt;
      decodeLeaf(r, first, cnt);
      if (uint64_t(first) + cnt > numTris) atomicAdd(bad, 1u);
    }
  }
}
__global__ void kValidateParents(const unsigned* refcnt, uint32_t numNodes, unsigned* bad) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numNodes) return;
  if (refcnt[i] != (i == 0 ? 0u : 1u)) atomicAdd(bad, 1u);
}

// ------------------------------------------------------ 4-wide collapse
// One frontier level per launch: each 4-wide node takes a binary node and
// repeatedly opens its largest-area internal child until it holds four
// children (or only leaves remain).  Internal children get the next node ids
// (siblings are allocated together, so a node's children are adjacent).
__device__ __forceinline__ float slotArea(const float* lo, const float* hi) {
  float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return dx * dy + dy * dz + dz * dx;
}

__global__ void kCollapse4(const float4* n2, const int2* front, uint32_t nFront, float4* n4, unsigned* n4Count,
                           int2* next, unsigned* nextCount) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nFront) return;
  int src = front[i].x, dst = front[i].y;
  int ref[4];
  float lo[4][3], hi[4][3];
  auto loadChildren = [&](int node, int s0, int s1) {
    float4 a = n2[size_t(node) * 4 + 0], b = n2[size_t(node) * 4 + 1], c = n2[size_t(node) * 4 + 2],
           d = n2[size_t(node) * 4 + 3];
    lo[s0][0] = a.x, hi[s0][0] = a.y, lo[s0][1] = a.z, hi[s0][1] = a.w, lo[s0][2] = c.x, hi[s0][2] = c.y;
    lo[s1][0] = b.x, hi[s1][0] = b.y, lo[s1][1] = b.z, hi[s1][1] = b.w, lo[s1][2] = c.z, hi[s1][2] = c.w;
    ref[s0] = __float_as_int(d.x);
    ref[s1] = __float_as_int(d.y);
  };
  loadChildren(src, 0, 1);
  int cnt = 2;
  while (cnt < 4) {
    int pick = -1;
    float best = -1.f;
    for (int k = 0; k < cnt; ++k) {
      if (ref[k] < 0 || ref[k] == kEmptyRef) continue;
      float a = slotArea(lo[k], hi[k]);
      if (a > best) best = a, pick = k;
    }
    if (pick < 0) break;
    loadChildren(ref[pick], pick, cnt);
    ++cnt;
  }
  int internal = 0;
  for (int k = 0; k < cnt; ++k) internal += (ref[k] >= 0 && ref[k] != kEmptyRef) ? 1 : 0;
  unsigned base = internal ? atomicAdd(n4Count, unsigned(internal)) : 0u;
  unsigned fbase = internal ? atomicAdd(nextCount, unsigned(internal)) : 0u;
  int out[4];
  const float far = 3.0e38f;
  for (int k = 0; k < 4; ++k) {
    if (k >= cnt) {
      out[k] = kEmptyRef;
      lo[k][0] = lo[k][1] = lo[k][2] = hi[k][0] = hi[k][1] = hi[k][2] = far;
    } else if (ref[k] >= 0 && ref[k] != kEmptyRef) {
      out[k] = int(base);
      next[fbase] = make_int2(ref[k], int(base));
      ++base;
      ++fbase;
    } else {
      out[k] = ref[k];
    }
  }
  float4* o = n4 + size_t(dst) * 8;
  o[0] = make_float4(lo[0][0], lo[1][0], lo[2][0], lo[3][0]);
  o[1] = make_float4(hi[0][0], hi[1][0], hi[2][0], hi[3][0]);
  o[2] = make_float4(lo[0][1], lo[1][1], lo[2][1], lo[3][1]);
  o[3] = make_float4(hi[0][1], hi[1][1], hi[2][1], hi[3][1]);
  o[4] = make_float4(lo[0][2], lo[1][2], lo[2][2], lo[3][2]);
  o[5] = make_float4(hi[0][2], hi[1][2], hi[2][2], hi[3][2]);
  o[6] = make_float4(__int_as_float(out[0]), __int_as_float(out[1]), __int_as_float(out[2]), __int_as_float(out[3]));
  o[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void kValidate4(const float4* nodes, uint32_t numNodes, uint32_t numTris, unsigned* refcnt, unsigned* bad) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numNodes) return;
  float4 r4 = nodes[size_t(i) * 8 + 6];
  int32_t refs[4] = {__float_as_int(r4.x), __float_as_int(r4.y), __float_as_int(r4.z), __float_as_int(r4.w)};
  for (int q = 0; q < 4; ++q) {
    int32_t r = refs[q];
    if (r == kEmptyRef) continue;
    if (r >= 0) {
      if (uint32_t(r) >= numNodes || r == 0) atomicAdd(bad, 1u);
      else atomicAdd(&refcnt[r], 1u);
    } else {
      uint32_t first, cnt;
      decodeLeaf(r, first, cnt);
      if (uint64_t(first) + cnt > numTris) atomicAdd(bad, 1u);
    }
  }
}

// Quantized copy of a 4-wide node (64 bytes): per axis the origin (the
// smallest child plane) and a power-of-two scale 2^e, the smallest that
// spans the largest child plane in 255 steps; each child plane as a byte,
// rounded outwards: the lower plane q is the largest with origin + q * 2^e
// <= plane, the upper the smallest with origin + q * 2^e >= plane, both
// evaluated in float exactly as the traversal evaluates them (the product
// is exact, the sum rounds once).  Layout: (origin.xyz, biased exponents
// in bytes 0..2), (lo.x, hi.x, lo.y, hi.y bytes of children 0..3),
// (lo.z, hi.z, ref0, ref1), (ref2, ref3, 0, 0).
__device__ __forceinline__ float pow2f(int e) { return __uint_as_float(uint32_t(e + 127) << 23); }
__global__ void kQuantize4(const float4* n4, uint32_t numNodes, float4* nq) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numNodes) return;
  const float4* src = n4 + size_t(i) * 8;
  float pl[3][2][4];  // [axis][lo/hi][child]
  for (int ax = 0; ax < 3; ++ax)
    for (int side = 0; side < 2; ++side) {
      float4 v = src[2 * ax + side];
      pl[ax][side][0] = v.x, pl[ax][side][1] = v.y, pl[ax][side][2] = v.z, pl[ax][side][3] = v.w;
    }
  float4 rf = src[6];
  int ref[4] = {__float_as_int(rf.x), __float_as_int(rf.y), __float_as_int(rf.z), __float_as_int(rf.w)};
  float origin[3];
  uint32_t qlo[3], qhi[3], ebits = 0;
  for (int ax = 0; ax < 3; ++ax) {
    float o = FLT_MAX, m = -FLT_MAX;
    bool any = false;
    for (int c = 0; c < 4; ++c)
      if (ref[c] != kEmptyRef) {
        o = fminf(o, pl[ax][0][c]);
        m = fmaxf(m, pl[ax][1][c]);
        any = true;
      }
    if (!any) o = m = 0.f;
    double ext = double(m) - double(o);
    int e = -126;
    if (ext > 0.0) e = max(-126, min(127, int(ceil(log2(ext / 255.0)))));
    uint32_t pk = 0, ph = 0;
    for (;; ++e) {
      const float sc = pow2f(e);
      bool ok = true;
      pk = ph = 0;
      for (int c = 0; c < 4 && ok; ++c) {
        if (ref[c] == kEmptyRef) continue;
        const float lo = pl[ax][0][c], hi = pl[ax][1][c];
        double fl = floor((double(lo) - double(o)) / double(sc));
        uint32_t ql = fl <= 0.0 ? 0u : (fl >= 255.0 ? 255u : uint32_t(fl));
        while (ql > 0 && o + float(ql) * sc > lo) --ql;
        double fh = ceil((double(hi) - double(o)) / double(sc));
        uint32_t qh = fh <= 0.0 ? 0u : (fh >= 256.0 ? 256u : uint32_t(fh));
        while (qh < 256u && o + float(qh) * sc < hi) ++qh;
        if (qh > 255u) {
          ok = false;
        } else {
          pk |= ql << (8 * c);
          ph |= qh << (8 * c);
        }
      }
      if (ok || e >= 127) break;
    }
    origin[ax] = o;
    qlo[ax] = pk;
    qhi[ax] = ph;
    ebits |= uint32_t(e + 127) << (8 * ax);
  }
  float4* d = nq + size_t(i) * 4;
  d[0] = make_float4(origin[0], origin[1], origin[2], __uint_as_float(ebits));
  d[1] = make_float4(__uint_as_float(qlo[0]), __uint_as_float(qhi[0]), __uint_as_float(qlo[1]), __uint_as_float(qhi[1]));
  d[2] = make_float4(__uint_as_float(qlo[2]), __uint_as_float(qhi[2]), rf.x, rf.y);
  d[3] = make_float4(rf.z, rf.w, 0.f, 0.f);
}

template <class T> struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// ------------------------------------------------------- scan and sort
// Exclusive prefix sum of uint32 values in three passes, deterministic:
// per-2048-element tile sums, one block scanning the tile sums, then each
// tile's scan plus its offset.  Work per thread: 8 consecutive elements.
constexpr uint32_t kScanTile = 2048;
__device__ __forceinline__ uint32_t blockExclusive(uint32_t v, uint32_t* sh, uint32_t& total) {
  // exclusive scan of one value per thread over a 256-thread block
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t x = v;
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t k = 0; k < w; ++k) base += sh[k];
  total = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return base + x - v;
}
__global__ __launch_bounds__(256) void kScanTileSums(const uint32_t* in, uint32_t n, uint32_t* tileSums) {
  __shared__ uint32_t sh[4];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 8;
  uint32_t s = 0;
  for (uint32_t k = 0; k < 8; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t total;
  (void)blockExclusive(s, sh, total);
  if (threadIdx.x == 0) tileSums[blockIdx.x] = total;
}
__global__ __launch_bounds__(256) void kScanOfSums(uint32_t* sums, uint32_t nt) {
  // one block: each thread scans a contiguous run of the tile sums
  __shared__ uint32_t sh[4];
  const uint32_t per = (nt + 255) / 256, b0 = threadIdx.x * per, b1 = min(nt, b0 + per);
  uint32_t s = 0;
  for (uint32_t k = b0; k < b1; ++k) s += sums[k];
  uint32_t total, run = blockExclusive(s, sh, total);
  for (uint32_t k = b0; k < b1; ++k) {
    const uint32_t v = sums[k];
    sums[k] = run;
    run += v;
  }
}
__global__ __launch_bounds__(256) void kScanApply(const uint32_t* in, uint32_t n, const uint32_t* tileOff,
                                                  uint32_t* out) {
  __shared__ uint32_t sh[4];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 8;
  uint32_t v[8], s = 0;
  for (uint32_t k = 0; k < 8; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t total, run = blockExclusive(s, sh, total) + tileOff[blockIdx.x];
  for (uint32_t k = 0; k < 8; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}
// Scratch words the scan of n values needs (the tile sums).
static size_t scanScratch(uint32_t n) { return std::max<size_t>(1, (size_t(n) + kScanTile - 1) / kScanTile); }
static void exclusiveScan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch, hipStream_t s) {
  if (!n) return;
  const uint32_t nt = uint32_t(scanScratch(n));
  kScanTileSums<<<nt, 256, 0, s>>>(in, n, scratch);
  kScanOfSums<<<1, 256, 0, s>>>(scratch, nt);
  kScanApply<<<nt, 256, 0, s>>>(in, n, scratch, out);
}

// Stable LSD radix sort of (64-bit key, 32-bit value) pairs, 8-bit digits.
// Per pass: each 2048-key tile counts its digits (written digit-major, so
// that one exclusive scan of the counts gives every (digit, tile) its
// output base), then scatters its keys in their input order: the tile is
// walked in 8 rows of 256 keys; within a row a lane's rank among the lanes
// of its wave with the same digit comes from 8 ballots, and the waves'
// per-digit counts (in LDS) order the waves, so equal digits keep their
// input order and the sort is stable.
constexpr uint32_t kSortTile = 2048, kRadix = 256;
__global__ __launch_bounds__(256) void kRadixCount(const uint64_t* keys, uint32_t n, uint32_t shift,
                                                   uint32_t nTiles, uint32_t* counts) {
  __shared__ uint32_t h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  for (uint32_t k = threadIdx.x; k < kSortTile; k += 256)
    if (base + k < n) atomicAdd(&h[uint32_t(keys[base + k] >> shift) & (kRadix - 1)], 1u);
  __syncthreads();
  counts[size_t(threadIdx.x) * nTiles + blockIdx.x] = h[threadIdx.x];
}
__global__ __launch_bounds__(256) void kRadixScatter(const uint64_t* keys, const uint32_t* vals, uint32_t n,
                                                     uint32_t shift, uint32_t nTiles, const uint32_t* offsets,
                                                     uint64_t* keysOut, uint32_t* valsOut) {
  __shared__ uint32_t run[kRadix];      // keys of each digit already placed by this tile
  __shared__ uint32_t wcnt[4][kRadix];  // per-wave digit counts of the current row
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  run[t] = offsets[size_t(t) * nTiles + blockIdx.x];
  for (uint32_t k = 0; k < 4; ++k) wcnt[k][t] = 0;
  __syncthreads();
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t base = blockIdx.x * kSortTile;
  for (uint32_t row = 0; row < kSortTile / 256; ++row) {
    const uint32_t i = base + row * 256 + t;
    const bool in = i < n;
    uint64_t key = 0;
    uint32_t val = 0, d = 0;
    if (in) {
      key = keys[i];
      val = vals[i];
      d = uint32_t(key >> shift) & (kRadix - 1);
    }
    uint64_t match = __ballot(in);
    for (uint32_t b = 0; b < 8; ++b) {
      const uint64_t m = __ballot(in && ((d >> b) & 1u));
      match &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = uint32_t(__popcll(match & below));
    if (in && rank == 0) wcnt[w][d] = uint32_t(__popcll(match));
    __syncthreads();
    if (in) {
      uint32_t pos = run[d] + rank;
      for (uint32_t k = 0; k < w; ++k) pos += wcnt[k][d];
      keysOut[pos] = key;
      valsOut[pos] = val;
    }
    __syncthreads();
    run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
    for (uint32_t k = 0; k < 4; ++k) wcnt[k][t] = 0;
    __syncthreads();
  }
}

// Sorts n (key, value) pairs by the low `bits` bits of the key, stably, in
// passes of 8 bits; each pass swaps the (keys, keysTmp) and (vals, valsTmp)
// pointers, so the result ends in keys / vals.  counts and offsets hold
// 256 x tiles words, scanScr scanScratch(256 x tiles) words.
static void radixSortPairs(uint64_t*& keys, uint64_t*& keysTmp, uint32_t*& vals, uint32_t*& valsTmp, uint32_t n,
                           uint32_t bits, uint32_t* counts, uint32_t* offsets, uint32_t* scanScr, hipStream_t s) {
  const uint32_t nTiles = (n + kSortTile - 1) / kSortTile;
  for (uint32_t shift = 0; shift < bits; shift += 8) {
    kRadixCount<<<nTiles, 256, 0, s>>>(keys, n, shift, nTiles, counts);
    exclusiveScan(counts, offsets, kRadix * nTiles, scanScr, s);
    kRadixScatter<<<nTiles, 256, 0, s>>>(keys, vals, n, shift, nTiles, offsets, keysTmp, valsTmp);
    std::swap(keys, keysTmp);
    std::swap(vals, valsTmp);
  }
}


}  // namespace

int buildBvh(const BuildArgs& a, hipStream_t s, BuildResult& r, char* err, size_t errLen) {
  hipEvent_t e0, e1;
  HIPB(hipEventCreate(&e0));
  HIPB(hipEventCreate(&e1));
  HIPB(hipEventRecord(e0, s));
  const uint32_t n = a.numTris;
  if (a.leafMax < 1 || a.leafMax > uint32_t(kLeafMax)) {
    std::snprintf(err, errLen, "bvh: leaf_max_tris must be in [1, %d]", kLeafMax);
    return ASH_E_INVALID;
  }
  if (n >= (1u << 28)) {
    std::snprintf(err, errLen, "bvh: %u triangles exceed the 2^28 leaf addressing limit", n);
    return ASH_E_UNSUPPORTED;
  }
  DevBuf<float4> wtri, boxLo, boxHi, partLo, partHi, nodeLo, nodeHi;
  DevBuf<uint64_t> keys, keys2;
  DevBuf<uint32_t> ids, ids2, emitted, newIdx;
  DevBuf<int> left, right, parent, leafParent;
  DevBuf<int2> range;
  DevBuf<unsigned> maxLevel, maxDepth;
  DevBuf<int> level;
  DevBuf<float> cost;
  DevBuf<uint8_t> collapsed;
  DevBuf<uint32_t> counts, countOffs, scanScr;
  size_t nn = n ? n : 1;
  HIPB(hipMalloc(&wtri.p, nn * 3 * sizeof(float4)));
  HIPB(hipMalloc(&boxLo.p, nn * sizeof(float4)));
  HIPB(hipMalloc(&boxHi.p, nn * sizeof(float4)));
  HIPB(hipMalloc(&keys.p, nn * sizeof(uint64_t)));
  HIPB(hipMalloc(&keys2.p, nn * sizeof(uint64_t)));
  HIPB(hipMalloc(&ids.p, nn * sizeof(uint32_t)));
  HIPB(hipMalloc(&ids2.p, nn * sizeof(uint32_t)));
  const uint32_t nPart = 512;
  HIPB(hipMalloc(&partLo.p, nPart * sizeof(float4)));
  HIPB(hipMalloc(&partHi.p, nPart * sizeof(float4)));
  HIPB(hipMalloc(&maxDepth.p, sizeof(unsigned)));
  HIPB(hipMemsetAsync(maxDepth.p, 0, sizeof(unsigned), s));
  unsigned grid = (n + kBlock - 1) / kBlock;
  if (n) {
    kWorldTris<<<grid, kBlock, 0, s>>>(a, wtri.p, boxLo.p, boxHi.p);
    kCentroidBounds<<<nPart, kBlock, 0, s>>>(boxLo.p, boxHi.p, n, partLo.p, partHi.p);
    kMorton<<<grid, kBlock, 0, s>>>(boxLo.p, boxHi.p, n, partLo.p, partHi.p, nPart, keys.p, ids.p);
    const uint32_t nTiles = (n + kSortTile - 1) / kSortTile, nCounts = kRadix * nTiles;
    HIPB(hipMalloc(&counts.p, size_t(nCounts) * sizeof(uint32_t)));
    HIPB(hipMalloc(&countOffs.p, size_t(nCounts) * sizeof(uint32_t)));
    HIPB(hipMalloc(&scanScr.p, std::max(scanScratch(nCounts), scanScratch(n)) * sizeof(uint32_t)));
    radixSortPairs(keys.p, keys2.p, ids.p, ids2.p, n, 63, counts.p, countOffs.p, scanScr.p, s);
    std::swap(keys.p, keys2.p);  // the sorted pairs live in keys2 / ids2 from here on
    std::swap(ids.p, ids2.p);
    HIPB(hipGetLastError());
  }
  const uint32_t* sorted = ids2.p;
  uint32_t nInternal = n > 1 ? n - 1 : 0;
  uint32_t numNodes = 1;
  uint32_t depth = 1;
  double sah = 0.0;
  bool singleLeaf = n <= 1;
  if (nInternal) {
    size_t ni = nInternal;
    HIPB(hipMalloc(&left.p, ni * sizeof(int)));
    HIPB(hipMalloc(&right.p, ni * sizeof(int)));
    HIPB(hipMalloc(&parent.p, ni * sizeof(int)));
    HIPB(hipMalloc(&leafParent.p, nn * sizeof(int)));
    HIPB(hipMalloc(&range.p, ni * sizeof(int2)));
    HIPB(hipMalloc(&level.p, ni * sizeof(int)));
    HIPB(hipMalloc(&maxLevel.p, sizeof(unsigned)));
    HIPB(hipMalloc(&nodeLo.p, ni * sizeof(float4)));
    HIPB(hipMalloc(&nodeHi.p, ni * sizeof(float4)));
    HIPB(hipMalloc(&cost.p, ni * sizeof(float)));
    HIPB(hipMalloc(&collapsed.p, ni));
    HIPB(hipMalloc(&emitted.p, ni * sizeof(uint32_t)));
    HIPB(hipMalloc(&newIdx.p, ni * sizeof(uint32_t)));
    HIPB(hipMemsetAsync(maxLevel.p, 0, sizeof(unsigned), s));
    unsigned gi = (nInternal + kBlock - 1) / kBlock;
    if (a.quality == 2) {
      // PLOC hierarchy over the Morton order, then depth-first leaf order
      DevBuf<int> cNode[2], nnb, leafParent2;
      DevBuf<float4> cLo[2], cHi[2];
      DevBuf<uint32_t> valid, offs, nodeCnt, posOf, sorted2;
      DevBuf<unsigned> idCounter;
      for (int b2 = 0; b2 < 2; ++b2) {
        HIPB(hipMalloc(&cNode[b2].p, nn * sizeof(int)));
        HIPB(hipMalloc(&cLo[b2].p, nn * sizeof(float4)));
        HIPB(hipMalloc(&cHi[b2].p, nn * sizeof(float4)));
      }
      HIPB(hipMalloc(&nnb.p, nn * sizeof(int)));
      HIPB(hipMalloc(&valid.p, nn * sizeof(uint32_t)));
      HIPB(hipMalloc(&offs.p, nn * sizeof(uint32_t)));
      HIPB(hipMalloc(&nodeCnt.p, ni * sizeof(uint32_t)));
      HIPB(hipMalloc(&posOf.p, nn * sizeof(uint32_t)));
      HIPB(hipMalloc(&sorted2.p, nn * sizeof(uint32_t)));
      HIPB(hipMalloc(&leafParent2.p, nn * sizeof(int)));
      HIPB(hipMalloc(&idCounter.p, sizeof(unsigned)));
      HIPB(hipMemsetAsync(idCounter.p, 0, sizeof(unsigned), s));
      int rootParent = -1;
      HIPB(hipMemcpyAsync(parent.p, &rootParent, sizeof(int), hipMemcpyHostToDevice, s));
      kPlocInit<<<grid, kBlock, 0, s>>>(n, ids2.p, boxLo.p, boxHi.p, cNode[0].p, cLo[0].p, cHi[0].p);
      uint32_t m = n, iters = 0;
      while (m > 1) {
        unsigned gm = (m + kBlock - 1) / kBlock;
        kPlocNearest<<<gm, kBlock, 0, s>>>(m, cLo[0].p, cHi[0].p, nnb.p);
        kPlocMerge<<<gm, kBlock, 0, s>>>(m, n, nnb.p, cNode[0].p, cLo[0].p, cHi[0].p, cNode[1].p, cLo[1].p, cHi[1].p,
                                          valid.p, left.p, right.p, parent.p, leafParent.p, nodeCnt.p, idCounter.p);
        exclusiveScan(valid.p, offs.p, m, scanScr.p, s);
        kPlocCompact<<<gm, kBlock, 0, s>>>(m, valid.p, offs.p, cNode[1].p, cLo[1].p, cHi[1].p, cNode[0].p, cLo[0].p,
                                            cHi[0].p);
        uint32_t lastOff = 0, lastValid = 0;
        HIPB(hipMemcpyAsync(&lastOff, offs.p + m - 1, 4, hipMemcpyDeviceToHost, s));
        HIPB(hipMemcpyAsync(&lastValid, valid.p + m - 1, 4, hipMemcpyDeviceToHost, s));
        HIPB(hipStreamSynchronize(s));
        uint32_t next = lastOff + lastValid;
        if (next >= m || ++iters > 100000) {
          std::snprintf(err, errLen, "bvh: clustering made no progress (%u clusters)", m);
          return ASH_E_INVALID;
        }
        m = next;
      }
      unsigned used = 0;
      HIPB(hipMemcpyAsync(&used, idCounter.p, sizeof used, hipMemcpyDeviceToHost, s));
      HIPB(hipStreamSynchronize(s));
      if (used != nInternal) {
        std::snprintf(err, errLen, "bvh: clustering produced %u of %u internal nodes", used, nInternal);
        return ASH_E_INVALID;
      }
      kPlocLeafOrder<<<grid, kBlock, 0, s>>>(n, left.p, right.p, parent.p, leafParent.p, nodeCnt.p, ids2.p, posOf.p,
                                             sorted2.p, leafParent2.p);
      kPlocRanges<<<gi, kBlock, 0, s>>>(nInternal, left.p, posOf.p, nodeCnt.p, range.p);
      kPlocRelabel<<<gi, kBlock, 0, s>>>(nInternal, left.p, right.p, posOf.p);
      HIPB(hipMemcpyAsync(ids2.p, sorted2.p, nn * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
      HIPB(hipMemcpyAsync(leafParent.p, leafParent2.p, nn * sizeof(int), hipMemcpyDeviceToDevice, s));
      HIPB(hipGetLastError());
    } else {
      kKarras<<<gi, kBlock, 0, s>>>(keys2.p, int(n), left.p, right.p, parent.p, leafParent.p, range.p);
    }
    kNodeLevel<<<gi, kBlock, 0, s>>>(int(nInternal), parent.p, level.p, maxLevel.p);
    unsigned deepest = 0;
    HIPB(hipMemcpyAsync(&deepest, maxLevel.p, sizeof deepest, hipMemcpyDeviceToHost, s));
    HIPB(hipStreamSynchronize(s));
    if (deepest >= nInternal) {
      std::snprintf(err, errLen, "bvh: tree deeper (%u) than its node count", deepest);
      return ASH_E_INVALID;
    }
    for (int lv = int(deepest); lv >= 0; --lv)
      kRefitLevel<<<gi, kBlock, 0, s>>>(int(nInternal), lv, level.p, sorted, boxLo.p, boxHi.p, left.p, right.p,
                                        range.p, nodeLo.p, nodeHi.p, cost.p, collapsed.p, a.leafMax, a.quality);
    kEmitFlags<<<gi, kBlock, 0, s>>>(int(nInternal), parent.p, collapsed.p, emitted.p, maxDepth.p);
    exclusiveScan(emitted.p, newIdx.p, nInternal, scanScr.p, s);
    uint32_t lastIdx = 0, lastEm = 0;
    uint8_t rootCollapsed = 0;
    float rootCost = 0.f;
    float4 rl, rh;
    HIPB(hipMemcpyAsync(&lastIdx, newIdx.p + nInternal - 1, 4, hipMemcpyDeviceToHost, s));
    HIPB(hipMemcpyAsync(&lastEm, emitted.p + nInternal - 1, 4, hipMemcpyDeviceToHost, s));
    HIPB(hipMemcpyAsync(&rootCollapsed, collapsed.p, 1, hipMemcpyDeviceToHost, s));
    HIPB(hipMemcpyAsync(&rootCost, cost.p, 4, hipMemcpyDeviceToHost, s));
    HIPB(hipMemcpyAsync(&rl, nodeLo.p, 16, hipMemcpyDeviceToHost, s));
    HIPB(hipMemcpyAsync(&rh, nodeHi.p, 16, hipMemcpyDeviceToHost, s));
    HIPB(hipMemcpyAsync(&depth, maxDepth.p, 4, hipMemcpyDeviceToHost, s));
    HIPB(hipStreamSynchronize(s));
    float ra = (rh.x - rl.x) * (rh.y - rl.y) + (rh.y - rl.y) * (rh.z - rl.z) + (rh.z - rl.z) * (rh.x - rl.x);
    sah = ra > 0 ? double(rootCost) / double(ra) : 0.0;
    if (rootCollapsed) {
      singleLeaf = true;
    } else {
      numNodes = lastIdx + lastEm;
      HIPB(hipMalloc(&r.nodes, size_t(numNodes) * 4 * sizeof(float4)));
      kEmitNodes<<<gi, kBlock, 0, s>>>(int(nInternal), emitted.p, newIdx.p, left.p, right.p, range.p, collapsed.p,
                                       sorted, boxLo.p, boxHi.p, nodeLo.p, nodeHi.p, r.nodes);
    }
  }
  if (singleLeaf) {
    numNodes = 1;
    depth = 1;
    HIPB(hipMalloc(&r.nodes, 4 * sizeof(float4)));
    kSingleLeafRoot<<<1, 1, 0, s>>>(n, boxLo.p, boxHi.p, sorted, r.nodes);
  }
  HIPB(hipMalloc(&r.tris, nn * 3 * sizeof(float4)));
  HIPB(hipMalloc(&r.shade, nn * 6 * sizeof(float4)));
  if (n) kReorderTris<<<grid, kBlock, 0, s>>>(n, sorted, wtri.p, r.tris);
  if (n) kShadeTris<<<grid, kBlock, 0, s>>>(n, sorted, a, wtri.p, r.shade);
  HIPB(hipGetLastError());
  {
    DevBuf<unsigned> refcnt, bad;
    HIPB(hipMalloc(&refcnt.p, size_t(numNodes) * sizeof(unsigned)));
    HIPB(hipMalloc(&bad.p, sizeof(unsigned)));
    HIPB(hipMemsetAsync(refcnt.p, 0, size_t(numNodes) * sizeof(unsigned), s));
    HIPB(hipMemsetAsync(bad.p, 0, sizeof(unsigned), s));
    unsigned gv = (numNodes + kBlock - 1) / kBlock;
    kValidate<<<gv, kBlock, 0, s>>>(r.nodes, numNodes, n, refcnt.p, bad.p);
    kValidateParents<<<gv, kBlock, 0, s>>>(refcnt.p, numNodes, bad.p);
    unsigned nbad = 0;
    HIPB(hipMemcpyAsync(&nbad, bad.p, sizeof nbad, hipMemcpyDeviceToHost, s));
    HIPB(hipStreamSynchronize(s));
    if (nbad) {
      std::snprintf(err, errLen, "bvh: structural validation failed (%u violations)", nbad);
      return ASH_E_INVALID;
    }
  }
  if (a.branching == 4) {
    DevBuf<int2> f0, f1;
    DevBuf<unsigned> cnts;  // [0] 4-wide nodes allocated, [1] next frontier size
    HIPB(hipMalloc(&r.nodes4, size_t(numNodes) * 8 * sizeof(float4)));
    HIPB(hipMalloc(&f0.p, size_t(numNodes) * sizeof(int2)));
    HIPB(hipMalloc(&f1.p, size_t(numNodes) * sizeof(int2)));
    HIPB(hipMalloc(&cnts.p, 2 * sizeof(unsigned)));
    int2 root = make_int2(0, 0);
    unsigned init[2] = {1u, 0u};
    HIPB(hipMemcpyAsync(f0.p, &root, sizeof root, hipMemcpyHostToDevice, s));
    HIPB(hipMemcpyAsync(cnts.p, init, sizeof init, hipMemcpyHostToDevice, s));
    uint32_t nFront = 1, levels = 0;
    while (nFront) {
      kCollapse4<<<(nFront + kBlock - 1) / kBlock, kBlock, 0, s>>>(r.nodes, f0.p, nFront, r.nodes4, cnts.p, f1.p,
                                                                    cnts.p + 1);
      unsigned c[2];
      HIPB(hipMemcpyAsync(c, cnts.p, sizeof c, hipMemcpyDeviceToHost, s));
      HIPB(hipMemsetAsync(cnts.p + 1, 0, sizeof(unsigned), s));
      HIPB(hipStreamSynchronize(s));
      std::swap(f0.p, f1.p);
      nFront = c[1];
      r.numNodes4 = c[0];
      ++levels;
      if (levels > 200 || r.numNodes4 > numNodes) {
        std::snprintf(err, errLen, "bvh: 4-wide collapse did not terminate");
        return ASH_E_INVALID;
      }
    }
    r.depth4 = levels;
    DevBuf<unsigned> refcnt, bad;
    HIPB(hipMalloc(&refcnt.p, size_t(r.numNodes4) * sizeof(unsigned)));
    HIPB(hipMalloc(&bad.p, sizeof(unsigned)));
    HIPB(hipMemsetAsync(refcnt.p, 0, size_t(r.numNodes4) * sizeof(unsigned), s));
    HIPB(hipMemsetAsync(bad.p, 0, sizeof(unsigned), s));
    unsigned gv = (r.numNodes4 + kBlock - 1) / kBlock;
    kValidate4<<<gv, kBlock, 0, s>>>(r.nodes4, r.numNodes4, n, refcnt.p, bad.p);
    kValidateParents<<<gv, kBlock, 0, s>>>(refcnt.p, r.numNodes4, bad.p);
    unsigned nbad = 0;
    HIPB(hipMemcpyAsync(&nbad, bad.p, sizeof nbad, hipMemcpyDeviceToHost, s));
    HIPB(hipStreamSynchronize(s));
    if (nbad) {
      std::snprintf(err, errLen, "bvh: 4-wide structural validation failed (%u violations)", nbad);
      return ASH_E_INVALID;
    }
    if (a.quantize) {
      HIPB(hipMalloc(&r.nodesQ, size_t(r.numNodes4) * 4 * sizeof(float4)));
      kQuantize4<<<gv, kBlock, 0, s>>>(r.nodes4, r.numNodes4, r.nodesQ);
      HIPB(hipGetLastError());
    }
  }
  HIPB(hipEventRecord(e1, s));
  HIPB(hipEventSynchronize(e1));
  float ms = 0;
  HIPB(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  r.numNodes = numNodes;
  r.maxDepth = depth;
  r.sahCost = sah;
  r.ms = ms;
  return ASH_OK;
}

}  // namespace ashd
