// Wavefront path tracing kernels for gfx950 (replaces raytrace.projective.rgen,
// the closest-hit shaders of the plugin set, the miss shaders and the
// hardware traversal of the Vulkan pipeline).
//
// One render batch = shots x frames x pixels paths.  Per bounce:
//   extend  : closest hit of every live path (persistent waves taking rays
//             from per-XCD queue slices; 4-wide BVH traversal, nearest child
//             first, with a per-thread stack of 12 entries in LDS spilling to
//             global memory; the camera rays walk the tree once per wave; the
//             binary tree remains selectable)
//   shade   : miss / light hit / material closest-hit logic; next rays and
//             shadow rays are appended to compacted queues with one atomic per
//             wave (ballot + popcount)
//   shadow  : any hit of the NEE rays; unoccluded paths add their pending Ld
// then accumulate folds the frames of each pixel in frame order into the
// running weighted mean exactly as the reference's per-frame image update.
#include "render.h"

namespace ashd {

namespace {

// ------------------------------------------------------------ wave utils
__device__ __forceinline__ int waveAppend(uint32_t* counter, bool pred) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return -1;
  int lane = __lane_id();
  int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
  base = __shfl(base, leader);
  uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
  return pred ? int(base + uint32_t(__popcll(below))) : -1;
}
__device__ __forceinline__ void waveCount64(unsigned long long* counter, bool pred) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  int lane = __lane_id();
  int leader = __ffsll((unsigned long long)mask) - 1;
  if (lane == leader) atomicAdd(counter, (unsigned long long)__popcll(mask));
}

// --------------------------------------------------------------- traversal
struct TRay {
  f3 o, d, inv;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

__device__ __forceinline__ float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ void setupRay(TRay& r, f3 o, f3 d) {
  r.o = o;
  r.d = d;
  r.inv = f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = r.kz + 1 == 3 ? 0 : r.kz + 1;
  r.ky = r.kx + 1 == 3 ? 0 : r.kx + 1;
  if (comp(d, r.kz) < 0.0f) {
    int t = r.kx;
    r.kx = r.ky;
    r.ky = t;
  }
  float dz = comp(d, r.kz);
  r.Sx = comp(d, r.kx) / dz;
  r.Sy = comp(d, r.ky) / dz;
  r.Sz = 1.0f / dz;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013) on the
// vertices relative to the ray origin, permuted to (kx, ky, kz).
__device__ __forceinline__ bool hitTriCore(const TRay& r, float Akx, float Aky, float Akz, float Bkx, float Bky,
                                           float Bkz, float Ckx, float Cky, float Ckz, float tmin, float tmax,
                                           float& t, float& u, float& v) {
  float Ax = Akx - r.Sx * Akz, Ay = Aky - r.Sy * Akz;
  float Bx = Bkx - r.Sx * Bkz, By = Bky - r.Sy * Bkz;
  float Cx = Ckx - r.Sx * Ckz, Cy = Cky - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
    V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
    W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  float T = U * Az + V * Bz + W * Cz;
  float rcp = 1.0f / det;
  float tt = T * rcp;
  if (!(tt >= tmin && tt <= tmax)) return false;
  t = tt;
  u = V * rcp;
  v = W * rcp;
  return true;
}
__device__ __forceinline__ bool hitTri(const TRay& r, float4 p0, float4 p1, float4 p2, float tmin, float tmax,
                                       float& t, float& u, float& v) {
  f3 A = f3{p0.x, p0.y, p0.z} - r.o, B = f3{p1.x, p1.y, p1.z} - r.o, C = f3{p2.x, p2.y, p2.z} - r.o;
  return hitTriCore(r, comp(A, r.kx), comp(A, r.ky), comp(A, r.kz), comp(B, r.kx), comp(B, r.ky), comp(B, r.kz),
                    comp(C, r.kx), comp(C, r.ky), comp(C, r.kz), tmin, tmax, t, u, v);
}

// Conservative slab test (far bound scaled by 1 + 2*gamma_3, Ize 2013
// "Robust BVH Ray Traversal").  Near and far planes are chosen by the sign
// of the inverse direction; an endpoint that is NaN (a ray parallel to a
// plane it starts on) is ignored by the min/max chains, so such a ray is
// never culled by that axis, whatever the sign of its zero component.
__device__ __forceinline__ bool slab(float lx, float hx, float ly, float hy, float lz, float hz, const TRay& r,
                                     float tmin, float tmax, float& tn) {
  float x0 = (lx - r.o.x) * r.inv.x, x1 = (hx - r.o.x) * r.inv.x;
  float y0 = (ly - r.o.y) * r.inv.y, y1 = (hy - r.o.y) * r.inv.y;
  float z0 = (lz - r.o.z) * r.inv.z, z1 = (hz - r.o.z) * r.inv.z;
  const bool sx = r.inv.x < 0.0f, sy = r.inv.y < 0.0f, sz = r.inv.z < 0.0f;
  float nx = sx ? x1 : x0, fx = sx ? x0 : x1;
  float ny = sy ? y1 : y0, fy = sy ? y0 : y1;
  float nz = sz ? z1 : z0, fz = sz ? z0 : z1;
  tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
  float tf = fminf(fminf(fx, fy), fminf(fz, tmax)) * 1.0000004f;
  return tn <= tf;
}

// Per-thread traversal stack: the first `nl` entries in LDS (stride = block
// size), then the rest in global memory (stride = all persistent threads),
// `cap` entries in all.  Entry distances are kept only by the 4-wide
// closest-hit walk.
struct Stack {
  int* lds;
  float* ldsT;
  int* spill;
  float* spillT;
  uint32_t stride;
  int nl;
  int cap;
  int sp;
  __device__ __forceinline__ void push(int v) {
    if (sp < nl) lds[sp * kTraceBlock] = v;
    else if (sp < cap) spill[size_t(sp - nl) * stride] = v;
    ++sp;
  }
  __device__ __forceinline__ int pop() {
    --sp;
    if (sp < nl) return lds[sp * kTraceBlock];
    if (sp < cap) return spill[size_t(sp - nl) * stride];
    return kEmptyRef;  // overflowed entry (the host sizes `cap` to the tree depth)
  }
  __device__ __forceinline__ void push(int v, float t) {
    if (sp < nl) {
      lds[sp * kTraceBlock] = v;
      ldsT[sp * kTraceBlock] = t;
    } else if (sp < cap) {
      spill[size_t(sp - nl) * stride] = v;
      spillT[size_t(sp - nl) * stride] = t;
    }
    ++sp;
  }
  __device__ __forceinline__ int pop(float& t) {
    --sp;
    if (sp < nl) {
      t = ldsT[sp * kTraceBlock];
      return lds[sp * kTraceBlock];
    }
    if (sp < cap) {
      t = spillT[size_t(sp - nl) * stride];
      return spill[size_t(sp - nl) * stride];
    }
    t = 0.f;
    return kEmptyRef;
  }
};

// ------------------------------------------------ 4-wide node traversal
// Tests the four child boxes of a 128-byte node and returns the hit
// children sorted by entry distance (sorting network; misses sort last).
struct Kids {
  float t[4];
  int c[4];
  int n;
};
__device__ __forceinline__ void cswap(Kids& k, int a, int b) {
  const bool sw = k.t[b] < k.t[a];
  const float ta = k.t[a], tb = k.t[b];
  const int ca = k.c[a], cb = k.c[b];
  k.t[a] = sw ? tb : ta;
  k.t[b] = sw ? ta : tb;
  k.c[a] = sw ? cb : ca;
  k.c[b] = sw ? ca : cb;
}
__device__ __forceinline__ void testNode4(const float4* np, const TRay& r, float tmin, float lim, Kids& k) {
  float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5], rf = np[6];
  const float kMiss = 3.0e38f;
  float t;
  bool h;
  h = slab(lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, r, tmin, lim, t);
  k.t[0] = h ? t : kMiss;
  k.n = h;
  h = slab(lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, r, tmin, lim, t);
  k.t[1] = h ? t : kMiss;
  k.n += h;
  h = slab(lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, r, tmin, lim, t);
  k.t[2] = h ? t : kMiss;
  k.n += h;
  h = slab(lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, r, tmin, lim, t);
  k.t[3] = h ? t : kMiss;
  k.n += h;
  k.c[0] = __float_as_int(rf.x), k.c[1] = __float_as_int(rf.y), k.c[2] = __float_as_int(rf.z),
  k.c[3] = __float_as_int(rf.w);
  cswap(k, 0, 1);
  cswap(k, 2, 3);
  cswap(k, 0, 2);
  cswap(k, 1, 3);
  cswap(k, 1, 2);
}

// The same test on a quantized 64-byte node: per-axis origin and power-of-
// two scale, 8-bit child planes rounded outwards by the builder (each
// dequantized box contains its exact box, computed with the expression
// used here), child refs, empty slots masked by their ref.  Because the
// slab arithmetic is monotone in the planes, a dequantized box is entered
// whenever its exact box is, so the triangles tested include every one the
// exact-box walks test.
__device__ __forceinline__ void testNode4Q(const float4* np, const TRay& r, float tmin, float lim, Kids& k) {
  const uint4* up = reinterpret_cast<const uint4*>(np);
  const float4 h0 = np[0];
  const uint4 q1 = up[1], q2 = up[2], q3 = up[3];
  const uint32_t e = __float_as_uint(h0.w);
  const float sx = __uint_as_float((e & 0xffu) << 23), sy = __uint_as_float(((e >> 8) & 0xffu) << 23),
              sz = __uint_as_float(((e >> 16) & 0xffu) << 23);
  k.c[0] = int(q2.z), k.c[1] = int(q2.w), k.c[2] = int(q3.x), k.c[3] = int(q3.y);
  const float kMiss = 3.0e38f;
  k.n = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t sh = 8u * uint32_t(c);
    float lx = h0.x + float((q1.x >> sh) & 0xffu) * sx, hx = h0.x + float((q1.y >> sh) & 0xffu) * sx;
    float ly = h0.y + float((q1.z >> sh) & 0xffu) * sy, hy = h0.y + float((q1.w >> sh) & 0xffu) * sy;
    float lz = h0.z + float((q2.x >> sh) & 0xffu) * sz, hz = h0.z + float((q2.y >> sh) & 0xffu) * sz;
    float t;
    bool h = slab(lx, hx, ly, hy, lz, hz, r, tmin, lim, t) && k.c[c] != kEmptyRef;
    k.t[c] = h ? t : kMiss;
    k.n += h;
  }
  cswap(k, 0, 1);
  cswap(k, 2, 3);
  cswap(k, 0, 2);
  cswap(k, 1, 3);
  cswap(k, 1, 2);
}

// Closest hit in [tmin, tmax]; ties at equal t go to the smaller global id.
// Returns the leaf position of the hit triangle or -1.
template <bool COUNT>
__device__ int traceClosest(const SceneView& sc, const TRay& r, float tmin, float tmax, Stack& st, float& bt,
                            float& bu, float& bv, uint32_t& nNodes, uint32_t& nTris) {
  int best = -1, bestGid = 0x7fffffff;
  bt = tmax;
  st.sp = 0;
  int node = 0;
  for (;;) {
    if (node >= 0) {
      if (node != kEmptyRef) {
        const float4* np = sc.nodes + size_t(node) * 4;
        float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
        if (COUNT) ++nNodes;
        float t0, t1;
        float lim = best >= 0 ? bt : tmax;
        bool h0 = slab(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, r, tmin, lim, t0);
        bool h1 = slab(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, r, tmin, lim, t1);
        int c0 = __float_as_int(n3.x), c1 = __float_as_int(n3.y);
        if (h0 && h1) {
          if (t1 < t0) {
            int tmp = c0;
            c0 = c1;
            c1 = tmp;
          }
          st.push(c1);
          node = c0;
          continue;
        }
        if (h0) {
          node = c0;
          continue;
        }
        if (h1) {
          node = c1;
          continue;
        }
      }
    } else {
      uint32_t first, cnt;
      decodeLeaf(node, first, cnt);
      for (uint32_t k = first; k < first + cnt; ++k) {
        const float4* tp = sc.tris + size_t(k) * 3;
        float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
        if (COUNT) ++nTris;
        float t, u, v;
        if (hitTri(r, p0, p1, p2, tmin, best >= 0 ? bt : tmax, t, u, v)) {
          int gid = __float_as_int(p2.w);
          if (best < 0 || t < bt || (t == bt && gid < bestGid)) {
            bt = t;
            bu = u;
            bv = v;
            best = int(k);
            bestGid = gid;
          }
        }
      }
    }
    if (st.sp == 0) break;
    node = st.pop();
  }
  return best;
}

template <bool COUNT>
__device__ bool traceAny(const SceneView& sc, const TRay& r, float tmin, float tmax, Stack& st, uint32_t& nNodes,
                         uint32_t& nTris) {
  st.sp = 0;
  int node = 0;
  for (;;) {
    if (node >= 0) {
      if (node != kEmptyRef) {
        const float4* np = sc.nodes + size_t(node) * 4;
        float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
        if (COUNT) ++nNodes;
        float t0, t1;
        bool h0 = slab(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, r, tmin, tmax, t0);
        bool h1 = slab(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, r, tmin, tmax, t1);
        int c0 = __float_as_int(n3.x), c1 = __float_as_int(n3.y);
        if (h0 && h1) {
          if (t1 < t0) {
            int tmp = c0;
            c0 = c1;
            c1 = tmp;
          }
          st.push(c1);
          node = c0;
          continue;
        }
        if (h0) {
          node = c0;
          continue;
        }
        if (h1) {
          node = c1;
          continue;
        }
      }
    } else {
      uint32_t first, cnt;
      decodeLeaf(node, first, cnt);
      for (uint32_t k = first; k < first + cnt; ++k) {
        const float4* tp = sc.tris + size_t(k) * 3;
        float t, u, v;
        if (COUNT) ++nTris;
        if (hitTri(r, tp[0], tp[1], tp[2], tmin, tmax, t, u, v)) return true;
      }
    }
    if (st.sp == 0) break;
    node = st.pop();
  }
  return false;
}

// 4-wide closest hit.  Children are visited nearest first; stacked entries
// keep their entry distance and are dropped on pop when it lies beyond the
// current closest hit.  The drop test is the slab test's own far clamp
// (tn <= lim * (1 + 2 gamma_3)), so the triangles tested include every one
// the binary walk tests and the result is identical (same tie rule).
// QUANT: walk the quantized 64-byte nodes instead of the 128-byte ones.
template <bool COUNT, bool QUANT>
__device__ int traceClosest4(const SceneView& sc, const TRay& r, float tmin, float tmax, Stack& st, float& bt,
                             float& bu, float& bv, uint32_t& nNodes, uint32_t& nTris) {
  int best = -1, bestGid = 0x7fffffff;
  bt = tmax;
  st.sp = 0;
  int node = 0;
  for (;;) {
    if (node >= 0) {
      if (node != kEmptyRef) {
        if (COUNT) ++nNodes;
        Kids k;
        if (QUANT) testNode4Q(sc.nodesQ + size_t(node) * 4, r, tmin, best >= 0 ? bt : tmax, k);
        else testNode4(sc.nodes4 + size_t(node) * 8, r, tmin, best >= 0 ? bt : tmax, k);
        if (k.n > 0) {
          if (k.n > 3) st.push(k.c[3], k.t[3]);
          if (k.n > 2) st.push(k.c[2], k.t[2]);
          if (k.n > 1) st.push(k.c[1], k.t[1]);
          node = k.c[0];
          continue;
        }
      }
    } else {
      uint32_t first, cnt;
      decodeLeaf(node, first, cnt);
      for (uint32_t k = first; k < first + cnt; ++k) {
        const float4* tp = sc.tris + size_t(k) * 3;
        float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
        if (COUNT) ++nTris;
        float t, u, v;
        if (hitTri(r, p0, p1, p2, tmin, best >= 0 ? bt : tmax, t, u, v)) {
          int gid = __float_as_int(p2.w);
          if (best < 0 || t < bt || (t == bt && gid < bestGid)) {
            bt = t;
            bu = u;
            bv = v;
            best = int(k);
            bestGid = gid;
          }
        }
      }
    }
    bool more = false;
    while (st.sp > 0) {
      float tn;
      int c = st.pop(tn);
      if (best < 0 || tn <= bt * 1.0000004f) {
        node = c;
        more = true;
        break;
      }
    }
    if (!more) break;
  }
  return best;
}

template <bool COUNT, bool QUANT>
__device__ bool traceAny4(const SceneView& sc, const TRay& r, float tmin, float tmax, Stack& st, uint32_t& nNodes,
                          uint32_t& nTris) {
  st.sp = 0;
  int node = 0;
  for (;;) {
    if (node >= 0) {
      if (node != kEmptyRef) {
        if (COUNT) ++nNodes;
        Kids k;
        if (QUANT) testNode4Q(sc.nodesQ + size_t(node) * 4, r, tmin, tmax, k);
        else testNode4(sc.nodes4 + size_t(node) * 8, r, tmin, tmax, k);
        if (k.n > 0) {
          if (k.n > 3) st.push(k.c[3]);
          if (k.n > 2) st.push(k.c[2]);
          if (k.n > 1) st.push(k.c[1]);
          node = k.c[0];
          continue;
        }
      }
    } else {
      uint32_t first, cnt;
      decodeLeaf(node, first, cnt);
      for (uint32_t k = first; k < first + cnt; ++k) {
        const float4* tp = sc.tris + size_t(k) * 3;
        float t, u, v;
        if (COUNT) ++nTris;
        if (hitTri(r, tp[0], tp[1], tp[2], tmin, tmax, t, u, v)) return true;
      }
    }
    if (st.sp == 0) break;
    node = st.pop();
  }
  return false;
}

// Pop for the closest-hit walks: the next stacked entry whose entry
// distance is not beyond the current hit (kDoneRef when none is left).
constexpr int32_t kDoneRef = -0x7fffffff - 1;  // walk finished (leaf refs stay above: first < 2^28)
constexpr int kShadowRefill = 48;  // idle lanes of a wave that trigger a shadow-ray refill
// A round's node descent ends once fewer than kLeafBreak of the wave's lanes
// still descending hold no postponed leaf: the last few would otherwise keep
// the whole wave stepping nodes for them (8 to 12 measured best on the 1080p
// workload; waiting for every lane, 10 % slower).
constexpr int kLeafBreak = 8;

__device__ __forceinline__ int popClosest(Stack& st, int best, float bt) {
  while (st.sp > 0) {
    float tn;
    int c = st.pop(tn);
    if (best < 0 || tn <= bt * 1.0000004f) return c;
  }
  return kDoneRef;
}

// ------------------------------------------------ two-level traversal
// The TLAS is a binary tree over the instances' world boxes, walked with
// the world ray.  At an instance leaf the ray enters the instance's BLAS in
// object space (origin and direction through the world-to-object
// transform; the direction is not normalised, so t is the same parameter
// in both spaces) and the walk stays there until the stack falls back to
// the depth it had on entry.  Culling (entry distance against the closest
// hit) and the tie rule (the smaller global triangle id at equal t) are
// those of the flat walks.

__device__ __forceinline__ void toObject(const DevInstance& in, const TRay& rw, TRay& ro) {
  const float* m = in.w2o;
  const f3 o = rw.o, d = rw.d;
  setupRay(ro,
           f3{m[0] * o.x + m[1] * o.y + m[2] * o.z + m[3], m[4] * o.x + m[5] * o.y + m[6] * o.z + m[7],
              m[8] * o.x + m[9] * o.y + m[10] * o.z + m[11]},
           f3{m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
              m[8] * d.x + m[9] * d.y + m[10] * d.z});
}

// A BLAS vertex in world space, rounded as the flat build rounds it
// (((m0*x + m1*y) + m2*z) + m3 per row, no contraction).
__device__ __forceinline__ float4 toWorld(const DevInstance& in, float4 p) {
  const float* m = in.o2w;
  return make_float4(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                     m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11], 0.f);
}

// One TLAS node step: pushes the farther of two entered children; returns
// the next node, or kDoneRef when neither child is entered.
__device__ __forceinline__ int tlasStep(const SceneView& sc, const TRay& rw, int node, float tmin, float lim,
                                        Stack& st) {
  const float4* np = sc.tlas + size_t(node) * 4;
  const float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
  float t0, t1;
  const bool h0 = slab(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, rw, tmin, lim, t0);
  const bool h1 = slab(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, rw, tmin, lim, t1);
  int c0 = __float_as_int(n3.x), c1 = __float_as_int(n3.y);
  if (h0 && h1) {
    if (t1 < t0) {
      const int c = c0;
      c0 = c1, c1 = c;
      const float t = t0;
      t0 = t1, t1 = t;
    }
    st.push(c1, t1);
    return c0;
  }
  return h0 ? c0 : (h1 ? c1 : kDoneRef);
}


// -------------------------------------------------------------- raygen
// raytrace.projective.rgen:26-90 for one (shot, frame, pixel).
__device__ __forceinline__ void cameraRay(const GpuCamera& cam, float px, float py, uint32_t& seed, f3& ro,
                                          f3& rd) {
  f3 origin = transformPoint(cam.cameraToWorld, mk1(0.f));
  if (cam.type == CameraTypePerspective) {
    ro = origin;
    f3 pCamera = transformPoint(cam.rasterToCamera, f3{px, py, 0.f});
    f3 r = makeNormal(pCamera);
    if (cam.aperture > 0.f) {
      float l0 = rnd(seed), l1 = rnd(seed);
      float dx, dy;
      concentricDisk(l0, l1, dx, dy);
      f3 pLens = f3{cam.aperture * dx, cam.aperture * dy, 0.f};
      float ft = cam.focalDistance / r.z;
      f3 pFocus = ft * r;
      ro = transformPoint(cam.cameraToWorld, pLens);
      r = pFocus - pLens;
    }
    rd = transformDirection(cam.cameraToWorld, r);
  } else {
    ro = origin;
    float rx = (px - cam.fxfycxcy[2]) / cam.fxfycxcy[0];
    float ry = (py - cam.fxfycxcy[3]) / cam.fxfycxcy[1];
    rd = transformDirection(cam.cameraToWorld, f3{rx, ry, 1.f});
  }
}

// Last unit whose first path (accumulate thread) is <= key.
template <bool ACC>
__device__ __forceinline__ uint32_t findUnit(const WorkUnit* u, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if ((ACC ? u[mid].accBase : u[mid].base) <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Pixel of position k of a film's pixel order: 8x8 tiles in row-major
// order over the largest multiple-of-8 rectangle, then the right strip and
// the bottom strip row by row.  Work units are runs of this order, so a
// wave of 64 consecutive paths covers one tile: its camera rays, and the
// bounces from where they land, start close together.
__device__ __forceinline__ uint32_t runPixel(uint32_t k, uint32_t W, uint32_t H) {
  const uint32_t Wt = W & ~7u, Ht = H & ~7u, nFull = Wt * Ht;
  if (k < nFull) {
    const uint32_t t = k >> 6, j = k & 63u, tw = Wt >> 3;
    return ((t / tw) * 8u + (j >> 3)) * W + (t % tw) * 8u + (j & 7u);
  }
  const uint32_t r = k - nFull, rw = W - Wt;
  if (r < rw * Ht) return (r / rw) * W + Wt + r % rw;
  return Ht * W + (r - rw * Ht);
}

__global__ __launch_bounds__(256) void kRaygen(BatchDesc b, Wavefront w) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= b.n) return;
  const WorkUnit un = b.units[findUnit<false>(b.units, b.nUnits, p)];
  uint32_t rel = p - un.base;
  uint32_t fslot = rel % un.nf;  // pixel-major: the unit's frames of one pixel are adjacent
  int32_t frame = un.f0 + int32_t(fslot);
  uint32_t shot = un.shot;
  const ShotParams& sp = b.shotTab[shot];
  const uint32_t pix = runPixel(un.pix0 + rel / un.nf, sp.width, sp.height);
  uint32_t x = pix % sp.width, y = pix / sp.width;
  uint32_t seed = xxhash32Seed(x, y, uint32_t(frame));
  float jx, jy;
  if (frame == 0) {
    jx = jy = 0.5f;
  } else {
    jx = rnd(seed);
    jy = rnd(seed);
  }
  f3 ro, rd;
  cameraRay(sp.cam, float(x) + jx, float(y) + jy, seed, ro, rd);
  // Gaussian filter weight, rgen:134-145 (exp(-8) folded to its correctly
  // rounded float value).
  const float expXY = 0.00033546262f;
  float ox = jx - 0.5f, oy = jy - 0.5f;
  float fw = fmaxf(0.0f, expf(-2.0f * ox * ox) - expXY) * fmaxf(0.0f, expf(-2.0f * oy * oy) - expXY);
  w.rayO[0][p] = f4(ro, 0.f);
  w.rayD[0][p] = f4(rd, __int_as_float(int(p)));
  // no throughput record: a depth-1 path's throughput is (1, 1, 1), pdf 0
  w.rad[p] = make_float4(0.f, 0.f, 0.f, fw);
  w.info[0][p] = make_uint4(seed, 1u, pix, shot | (frame == 0 ? 0x10000u : 0u));
  if (frame == 0) {
    size_t plane = size_t(sp.width) * sp.height;
    for (uint32_t c = 0; c < sp.pc.nMultiChannel; ++c) sp.images[(c + 1) * plane + pix] = make_float4(0.f, 0.f, 0.f, 1.f);
  }
  if (p == 0) w.ctr[0] = b.n;
}

// ---------------------------------------------------------------- extend
template <bool WIDE, int NL = kStackLds>
__device__ __forceinline__ void initStack(Stack& st, int* lds, float* ldsT, const Wavefront& w) {
  uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x;
  st.lds = lds + threadIdx.x;
  st.ldsT = WIDE ? ldsT + threadIdx.x : nullptr;
  st.spill = w.spill + gt;
  st.spillT = WIDE ? w.spillT + gt : nullptr;
  st.stride = w.spillThreads;
  st.nl = NL;
  st.cap = NL + int(w.spillCap);
}

// Work split of the persistent traversal kernels: a wave takes 64
// consecutive queue entries per atomic fetch.  XCD: the queue is cut into
// one contiguous slice per XCD; a wave drains the slice of its own XCD
// first, then helps with the others, so each XCD's L2 serves one run of the
// pixel-ordered queue instead of rays from the whole image.
struct ChunkFetch {
  uint32_t* ctr;  // first fetch counter (queue order) or first slice counter (XCD)
  uint32_t count;
  uint32_t slice, home, k;
  uint32_t lo, hi;
  __device__ __forceinline__ void slicesFor(uint32_t s) {
    lo = s * slice;
    hi = lo < count ? min(count, lo + slice) : lo;
  }
};
template <bool XCD>
__device__ __forceinline__ void chunkInit(ChunkFetch& f, uint32_t* ctr, uint32_t count) {
  f.ctr = ctr;
  f.count = count;
  f.k = 0;
  if (XCD) {
    f.slice = (count + kXcds - 1) / kXcds;
    f.home = blockIdx.x % kXcds;
    f.slicesFor(f.home);
  } else {
    f.slice = count;
    f.home = 0;
    f.lo = 0;
    f.hi = count;
  }
}
// Next chunk start for the calling wave (all lanes), CH queue entries per
// atomic fetch; false when the queue is drained.
template <bool XCD, uint32_t CH = 64>
__device__ __forceinline__ bool chunkNext(ChunkFetch& f, int lane, uint32_t& start) {
  for (;;) {
    uint32_t s = XCD ? (f.home + f.k) % kXcds : 0u;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(f.ctr + (XCD ? s * 32u : 0u), CH);
    base = __shfl(base, 0);
    if (base < f.hi - f.lo) {
      start = f.lo + base;
      return true;
    }
    if (!XCD || ++f.k == kXcds) return false;
    f.slicesFor((f.home + f.k) % kXcds);
  }
}

// TREE: 0 binary, 1 4-wide, 2 quantized 4-wide.
template <bool COUNT, int TREE, bool XCD, int NL>
__global__ __launch_bounds__(kTraceBlock) void kExtend(SceneView sc, Wavefront w, int cur) {
  __shared__ int lds[NL * kTraceBlock];
  __shared__ float ldsT[TREE ? NL * kTraceBlock : 1];
  Stack st;
  initStack<TREE != 0, NL>(st, lds, ldsT, w);
  const uint32_t count = w.ctr[cur];
  const float4 *qo = w.rayO[cur], *qd = w.rayD[cur];
  int lane = __lane_id();
  uint32_t nNodes = 0, nTris = 0;
  ChunkFetch f;
  chunkInit<XCD>(f, XCD ? &w.ctr[kXcdCtr] : &w.ctr[3], count);
  uint32_t start;
  while (chunkNext<XCD>(f, lane, start)) {
    uint32_t i = start + uint32_t(lane);
    if (i < f.hi) {
      TRay r;
      setupRay(r, xyz(qo[i]), xyz(qd[i]));
      float t = 0.f, u = 0.f, v = 0.f;
      int pos = TREE ? traceClosest4<COUNT, TREE == 2>(sc, r, kMinimum, kInf, st, t, u, v, nNodes, nTris)
                     : traceClosest<COUNT>(sc, r, kMinimum, kInf, st, t, u, v, nNodes, nTris);
      if (pos >= 0) {
        const float4* tp = sc.tris + size_t(pos) * 3;
        w.hitF[i] = make_float4(t, u, v, 0.f);
        w.hitI[i] = make_int4(__float_as_int(tp[0].w), __float_as_int(tp[1].w), __float_as_int(tp[2].w), pos);
      } else {
        w.hitI[i] = make_int4(-1, -1, -1, -1);
      }
    }
  }
  if (COUNT) {
    atomicAdd(&w.cnt64[0], (unsigned long long)nNodes);
    atomicAdd(&w.cnt64[1], (unsigned long long)nTris);
  }
}

// Persistent closest-hit walk with dynamic ray fetch (4-wide float nodes).
// Each lane walks its own ray in rounds of the speculative while-while form
// (Aila and Laine 2009): it descends through inner nodes until it stands at
// a leaf; a lane that reaches a leaf while holding none keeps it aside and
// descends further until every lane of the wave holds one; then the wave
// tests the leaves together, so node visits and triangle tests are not
// interleaved in one loop iteration.  Idle lanes take new rays (from the
// XCD slices, as the wave fetch does) as soon as REFILL of them are idle,
// so a wave does not wait for its slowest ray.  Visit order, culling and
// tie rule are those of traceClosest4 (speculation only delays the
// culling), so the closest hit is the same.
struct DynState {
  int node, post, best, bestGid;
  float bt, bu, bv;
};

template <bool COUNT>
__device__ __forceinline__ void dynLeaf(const SceneView& sc, const TRay& r, DynState& s, int ref, uint32_t& nT) {
  uint32_t first, cnt;
  decodeLeaf(ref, first, cnt);
  for (uint32_t k = first; k < first + cnt; ++k) {
    const float4* tp = sc.tris + size_t(k) * 3;
    float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
    if (COUNT) ++nT;
    float t, u, v;
    if (hitTri(r, p0, p1, p2, kMinimum, s.best >= 0 ? s.bt : kInf, t, u, v)) {
      int gid = __float_as_int(p2.w);
      if (s.best < 0 || t < s.bt || (t == s.bt && gid < s.bestGid)) {
        s.bt = t;
        s.bu = u;
        s.bv = v;
        s.best = int(k);
        s.bestGid = gid;
      }
    }
  }
}

// One descend-and-test round; true when the ray has finished.
template <bool COUNT>
__device__ __forceinline__ bool dynRound(const SceneView& sc, const TRay& r, DynState& s, Stack& st, uint32_t& nN,
                                         uint32_t& nT) {
  while (s.node >= 0) {
    if (s.node != kEmptyRef) {
      if (COUNT) ++nN;
      Kids k;
      testNode4(sc.nodes4 + size_t(s.node) * 8, r, kMinimum, s.best >= 0 ? s.bt : kInf, k);
      if (k.n > 0) {
        if (k.n > 3) st.push(k.c[3], k.t[3]);
        if (k.n > 2) st.push(k.c[2], k.t[2]);
        if (k.n > 1) st.push(k.c[1], k.t[1]);
        s.node = k.c[0];
      } else {
        s.node = popClosest(st, s.best, s.bt);
      }
    } else {
      s.node = popClosest(st, s.best, s.bt);
    }
    if (s.node < 0 && s.node != kDoneRef && s.post == kDoneRef) {
      s.post = s.node;
      s.node = popClosest(st, s.best, s.bt);
    }
    if (__popcll(__ballot(s.post == kDoneRef)) < kLeafBreak) break;
  }
  if (s.post != kDoneRef) {
    dynLeaf<COUNT>(sc, r, s, s.post, nT);
    s.post = kDoneRef;
  }
  if (s.node < 0 && s.node != kDoneRef) {
    dynLeaf<COUNT>(sc, r, s, s.node, nT);
    s.node = popClosest(st, s.best, s.bt);
  }
  return s.node == kDoneRef && s.post == kDoneRef;
}

// Hands the next queue entries of the wave's slices to its idle lanes (all
// lanes call it); `drained` once every slice has been handed out.
__device__ __forceinline__ bool fetchIdle(ChunkFetch& f, bool idle, int lane, bool& drained, uint32_t& idx) {
  const uint64_t m = __ballot(idle);
  const uint32_t n = uint32_t(__popcll(m));
  if (n == 0 || drained) return false;
  const uint32_t rank = uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
  for (;;) {
    const uint32_t sl = (f.home + f.k) % kXcds;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(f.ctr + sl * 32u, n);
    base = __shfl(base, 0);
    const uint32_t len = f.hi - f.lo;
    if (base < len) {
      idx = f.lo + base + rank;
      bool got = idle && idx < f.hi;
      if (base + n >= len) {
        if (++f.k == kXcds) drained = true;
        else f.slicesFor((f.home + f.k) % kXcds);
      }
      return got;
    }
    if (++f.k == kXcds) {
      drained = true;
      return false;
    }
    f.slicesFor((f.home + f.k) % kXcds);
  }
}

// Persistent closest-hit loop over queue `cur` (per-lane dynamic fetch).
template <bool COUNT, int REFILL>
__device__ __forceinline__ void extendDynLoop(const SceneView& sc, const Wavefront& w, int cur, Stack& st, int lane,
                                              uint32_t& nNodes, uint32_t& nTris) {
  const uint32_t count = w.ctr[cur];
  const float4 *qo = w.rayO[cur], *qd = w.rayD[cur];
  ChunkFetch f;
  chunkInit<true>(f, &w.ctr[kXcdCtr], count);
  bool active = false, drained = false;
  uint32_t slot = 0;
  TRay r;
  DynState s;
  for (;;) {
    const uint32_t nIdle = uint32_t(__popcll(__ballot(!active)));
    if (!drained && nIdle >= uint32_t(REFILL)) {
      uint32_t i = 0;
      if (fetchIdle(f, !active, lane, drained, i)) {
        slot = i;
        setupRay(r, xyz(qo[i]), xyz(qd[i]));
        s.node = 0;
        s.post = kDoneRef;
        s.best = -1;
        s.bestGid = 0x7fffffff;
        s.bt = kInf;
        s.bu = s.bv = 0.f;
        st.sp = 0;
        active = true;
      }
    }
    if (__ballot(active) == 0) {
      if (drained) break;
      continue;
    }
    if (active && dynRound<COUNT>(sc, r, s, st, nNodes, nTris)) {
      active = false;
      if (s.best >= 0) {
        const float4* tp = sc.tris + size_t(s.best) * 3;
        w.hitF[slot] = make_float4(s.bt, s.bu, s.bv, 0.f);
        w.hitI[slot] = make_int4(__float_as_int(tp[0].w), __float_as_int(tp[1].w), __float_as_int(tp[2].w), s.best);
      } else {
        w.hitI[slot] = make_int4(-1, -1, -1, -1);
      }
    }
  }
}

template <bool COUNT, int NL, int REFILL>
__global__ __launch_bounds__(kTraceBlock) void kExtendDyn(SceneView sc, Wavefront w, int cur) {
  __shared__ int lds[NL * kTraceBlock];
  __shared__ float ldsT[NL * kTraceBlock];
  Stack st;
  initStack<true, NL>(st, lds, ldsT, w);
  uint32_t nNodes = 0, nTris = 0;
  extendDynLoop<COUNT, REFILL>(sc, w, cur, st, __lane_id(), nNodes, nTris);
  if (COUNT) {
    atomicAdd(&w.cnt64[0], (unsigned long long)nNodes);
    atomicAdd(&w.cnt64[1], (unsigned long long)nTris);
  }
}

// ------------------------------------------------- coherent (camera) walk
// Closest hit of the camera-ray queue, one walk per wave.  A wave's 64
// rays are the frames of a few neighbouring pixels and nearly coincide, so
// the wave visits nodes together: each lane tests its own ray against the
// node's four boxes, a child is taken when any lane's ray enters it, and the
// wave stack keeps the node ref once plus every lane's entry distance, so a
// popped entry is visited only when some lane's ray still needs it (the
// per-ray walk's pop test, lane by lane).  Children are visited in the
// first lane's order.  Every lane tests every node and triangle the wave
// visits, with its own culling bound, so it visits a superset of what its
// own walk would; with the tie rule of traceClosest4 (equal t: smaller
// global id) the closest hit does not depend on the visit order and is the
// per-ray walk's.  Stack: kPacketLds entries per wave in LDS, then the
// per-thread global spill (entries of the wave's ref written by every lane),
// as deep as the per-ray stacks (three entries per 4-wide level).
//
// When every ray of the wave has the same direction signs and the same
// dominant axis (UNI, nearly always for camera rays), node planes and
// triangle vertices come through the scalar cache at wave-uniform offsets:
// the near and far plane quads of each axis, and the vertex components in
// the rays' (kx, ky, kz) order, so that neither the slab test's plane choice
// nor the triangle test's axis permutation costs a per-lane select.  The
// arithmetic is that of slab and hitTri, so the results are bit-identical.
// Other waves run the same walk with slab and hitTri as they are.
constexpr int kPacketLds = 16;

typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v8f __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
// Scalar loads at wave-uniform addresses; each block waits for its data,
// and its outputs are early-clobber (no output shares a register with the
// address or offsets that later loads of the block read).
// 48-byte triangle record.
__device__ __forceinline__ void loadTriS(const float4* tp, float4& p0, float4& p1, float4& p2) {
  v8f a;
  v4f b;
  asm volatile("s_load_dwordx8 %0, %2, 0x0\n\ts_load_dwordx4 %1, %2, 0x20\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(a), "=&s"(b)
               : "s"(tp)
               : "memory");
  p0 = make_float4(a[0], a[1], a[2], a[3]);
  p1 = make_float4(a[4], a[5], a[6], a[7]);
  p2 = make_float4(b[0], b[1], b[2], b[3]);
}

// Wave-uniform traits of the rays of a UNI packet: byte offsets, within a
// node, of the near and far plane quads of each axis (by the direction's
// sign: lo planes at 0 / 32 / 64, hi planes 16 bytes further), and of the
// kx, ky, kz components within a vertex.
struct PacketUni {
  uint32_t offN[3], offF[3];
  uint32_t offK[3];
  float okx, oky, okz;  // per lane: the origin's kx, ky, kz components
};
// The near and far plane quads of a node, and its child refs.
__device__ __forceinline__ void loadNodeNF(const float4* np, const PacketUni& q, v4f& nx, v4f& fx, v4f& ny, v4f& fy,
                                           v4f& nz, v4f& fz, v4f& rf) {
  asm volatile(
      "s_load_dwordx4 %0, %7, %8\n\t"
      "s_load_dwordx4 %1, %7, %9\n\t"
      "s_load_dwordx4 %2, %7, %10\n\t"
      "s_load_dwordx4 %3, %7, %11\n\t"
      "s_load_dwordx4 %4, %7, %12\n\t"
      "s_load_dwordx4 %5, %7, %13\n\t"
      "s_load_dwordx4 %6, %7, 0x60\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&s"(nx), "=&s"(fx), "=&s"(ny), "=&s"(fy), "=&s"(nz), "=&s"(fz), "=&s"(rf)
      : "s"(np), "s"(q.offN[0]), "s"(q.offF[0]), "s"(q.offN[1]), "s"(q.offF[1]), "s"(q.offN[2]), "s"(q.offF[2])
      : "memory");
}
// A triangle's vertex components in (kx, ky, kz) order, and its global id.
__device__ __forceinline__ void loadTriK(const float4* tp, const PacketUni& q, float (&a)[9], int& gid) {
  const float4 *b1 = tp + 1, *b2 = tp + 2;
  asm volatile(
      "s_load_dword %0, %10, %13\n\t"
      "s_load_dword %1, %10, %14\n\t"
      "s_load_dword %2, %10, %15\n\t"
      "s_load_dword %3, %11, %13\n\t"
      "s_load_dword %4, %11, %14\n\t"
      "s_load_dword %5, %11, %15\n\t"
      "s_load_dword %6, %12, %13\n\t"
      "s_load_dword %7, %12, %14\n\t"
      "s_load_dword %8, %12, %15\n\t"
      "s_load_dword %9, %12, 0xc\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&s"(a[0]), "=&s"(a[1]), "=&s"(a[2]), "=&s"(a[3]), "=&s"(a[4]), "=&s"(a[5]), "=&s"(a[6]), "=&s"(a[7]),
        "=&s"(a[8]), "=&s"(gid)
      : "s"(tp), "s"(b1), "s"(b2), "s"(q.offK[0]), "s"(q.offK[1]), "s"(q.offK[2])
      : "memory");
}

__device__ __forceinline__ void pswap(uint32_t (&key)[4], int (&ref)[4], float (&t)[4], int a, int b) {
  const bool sw = key[b] < key[a];  // wave-uniform
  const uint32_t ka = key[a], kb = key[b];
  const int ra = ref[a], rb = ref[b];
  const float ta = t[a], tb = t[b];
  key[a] = sw ? kb : ka;
  key[b] = sw ? ka : kb;
  ref[a] = sw ? rb : ra;
  ref[b] = sw ? ra : rb;
  t[a] = sw ? tb : ta;
  t[b] = sw ? ta : tb;
}

// slab() on planes already ordered near / far for the ray.
__device__ __forceinline__ bool slabNF(float nxp, float fxp, float nyp, float fyp, float nzp, float fzp,
                                       const TRay& r, float tmin, float tmax, float& tn) {
  const float nx = (nxp - r.o.x) * r.inv.x, fx = (fxp - r.o.x) * r.inv.x;
  const float ny = (nyp - r.o.y) * r.inv.y, fy = (fyp - r.o.y) * r.inv.y;
  const float nz = (nzp - r.o.z) * r.inv.z, fz = (fzp - r.o.z) * r.inv.z;
  tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
  const float tf = fminf(fminf(fx, fy), fminf(fz, tmax)) * 1.0000004f;
  return tn <= tf;
}

template <bool COUNT, bool UNI>
__device__ __forceinline__ void packetWalk(const SceneView& sc, const Wavefront& w, const TRay& r, const PacketUni& q,
                                           bool valid, int lane, uint32_t gt, int cap, int* sRef, float* sT,
                                           int& best, float& bt, float& bu, float& bv, uint32_t& nN, uint32_t& nT) {
  const float kMiss = 3.0e38f;
  float lim = valid ? kInf : -1.0f;  // culling bound: kInf, then the closest hit's t (-1: no ray)
  int bestGid = 0x7fffffff;
  best = -1;
  bt = kInf;
  bu = bv = 0.f;
  int node = 0, sp = 0;
  for (;;) {
    if (node >= 0) {
      if (COUNT && valid) ++nN;
      float tc[4];
      int rc[4];
      float t;
      if (UNI) {
        v4f nx, fx, ny, fy, nz, fz, rf;
        loadNodeNF(sc.nodes4 + size_t(node) * 8, q, nx, fx, ny, fy, nz, fz, rf);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          tc[c] = slabNF(nx[c], fx[c], ny[c], fy[c], nz[c], fz[c], r, kMinimum, lim, t) ? t : kMiss;
          rc[c] = __float_as_int(rf[c]);
        }
      } else {
        const float4* nd = sc.nodes4 + size_t(node) * 8;
        const float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5], rf = nd[6];
        tc[0] = slab(lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, r, kMinimum, lim, t) ? t : kMiss;
        tc[1] = slab(lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, r, kMinimum, lim, t) ? t : kMiss;
        tc[2] = slab(lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, r, kMinimum, lim, t) ? t : kMiss;
        tc[3] = slab(lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, r, kMinimum, lim, t) ? t : kMiss;
        rc[0] = __float_as_int(rf.x), rc[1] = __float_as_int(rf.y), rc[2] = __float_as_int(rf.z),
        rc[3] = __float_as_int(rf.w);
      }
      uint32_t key[4];
      int n = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const bool any = __ballot(tc[c] < kMiss) != 0 && rc[c] != kEmptyRef;
        n += any;
        // entry distance of the first lane (positive floats order as their bits); unused last
        key[c] = any ? __builtin_amdgcn_readfirstlane(__float_as_uint(tc[c])) : 0xffffffffu;
      }
      pswap(key, rc, tc, 0, 1);
      pswap(key, rc, tc, 2, 3);
      pswap(key, rc, tc, 0, 2);
      pswap(key, rc, tc, 1, 3);
      pswap(key, rc, tc, 1, 2);
      if (n > 0) {
#pragma unroll
        for (int j = 3; j >= 1; --j) {
          if (j < n) {
            if (sp < kPacketLds) {
              if (lane == 0) sRef[sp] = rc[j];
              sT[sp * 64 + lane] = tc[j];
            } else if (sp < cap) {
              const size_t o = size_t(sp - kPacketLds) * w.spillThreads + gt;
              w.spill[o] = rc[j];
              w.spillT[o] = tc[j];
            }
            ++sp;
          }
        }
        node = rc[0];
        continue;
      }
    } else {
      uint32_t first, cnt;
      decodeLeaf(node, first, cnt);
      for (uint32_t k = first; k < first + cnt; ++k) {
        if (COUNT && valid) ++nT;
        float t, u, v;
        bool h;
        int gid;
        if (UNI) {
          float a[9];
          loadTriK(sc.tris + size_t(k) * 3, q, a, gid);
          h = hitTriCore(r, a[0] - q.okx, a[1] - q.oky, a[2] - q.okz, a[3] - q.okx, a[4] - q.oky, a[5] - q.okz,
                         a[6] - q.okx, a[7] - q.oky, a[8] - q.okz, kMinimum, lim, t, u, v);
        } else {
          float4 p0, p1, p2;
          loadTriS(sc.tris + size_t(k) * 3, p0, p1, p2);
          h = hitTri(r, p0, p1, p2, kMinimum, lim, t, u, v);
          gid = __float_as_int(p2.w);
        }
        if (h) {
          if (best < 0 || t < bt || (t == bt && gid < bestGid)) {
            bt = t;
            bu = u;
            bv = v;
            best = int(k);
            bestGid = gid;
            lim = t;
          }
        }
      }
    }
    // pop: the next entry some lane's ray still enters before its closest hit
    node = kDoneRef;
    while (sp > 0) {
      --sp;
      int ref = kEmptyRef;
      float tl = kMiss;
      if (sp < kPacketLds) {
        ref = sRef[sp];
        tl = sT[sp * 64 + lane];
      } else if (sp < cap) {
        const size_t o = size_t(sp - kPacketLds) * w.spillThreads + gt;
        ref = w.spill[o];
        tl = w.spillT[o];
      }
      ref = __builtin_amdgcn_readfirstlane(ref);
      if (ref != kEmptyRef && __ballot(tl < kMiss && tl <= lim * 1.0000004f) != 0) {
        node = ref;
        break;
      }
    }
    if (node == kDoneRef) break;
  }
}

// 8 waves per SIMD (the scalar registers the walk holds would allow 7; the
// spills that fit it in 78 cost less than the lost wave: 3.13 against 3.35
// ms for the 1080p camera rays).  FETCH: rays per atomic fetch (a multiple
// of 64), walked as consecutive 64-ray packets.
template <bool COUNT, uint32_t FETCH>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void kExtendPacket(SceneView sc, Wavefront w, int cur) {
  __shared__ int sRef[kTraceBlock / 64][kPacketLds];
  __shared__ float sT[kTraceBlock / 64][kPacketLds * 64];
  const int lane = __lane_id(), wv = int(threadIdx.x) >> 6;
  const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int cap = kPacketLds + int(w.spillCap);
  const uint32_t count = w.ctr[cur];
  const float4 *qo = w.rayO[cur], *qd = w.rayD[cur];
  uint32_t nN = 0, nT = 0;
  ChunkFetch f;
  chunkInit<true>(f, &w.ctr[kXcdCtr], count);
  uint32_t start;
  while (chunkNext<true, FETCH>(f, lane, start)) {
   for (uint32_t sub = start; sub < start + FETCH && sub < f.hi; sub += 64) {
    const uint32_t i = sub + uint32_t(lane);
    const bool valid = i < f.hi;  // lane 0 always holds a ray
    TRay r;
    setupRay(r, valid ? xyz(qo[i]) : f3{0.f, 0.f, 0.f}, valid ? xyz(qd[i]) : f3{1.f, 1.f, 1.f});
    // UNI when the wave's rays share their direction signs and axis order
    // (the rays of invalid lanes take the wave's traits: their bound -1
    // rejects every box and triangle)
    const bool sx = __builtin_amdgcn_readfirstlane(int(r.inv.x < 0.0f)) != 0;
    const bool sy = __builtin_amdgcn_readfirstlane(int(r.inv.y < 0.0f)) != 0;
    const bool sz = __builtin_amdgcn_readfirstlane(int(r.inv.z < 0.0f)) != 0;
    const int kx = __builtin_amdgcn_readfirstlane(r.kx), ky = __builtin_amdgcn_readfirstlane(r.ky),
              kz = __builtin_amdgcn_readfirstlane(r.kz);
    const bool same = !valid || ((r.inv.x < 0.0f) == sx && (r.inv.y < 0.0f) == sy && (r.inv.z < 0.0f) == sz &&
                                 r.kx == kx && r.ky == ky && r.kz == kz);
    PacketUni q;
    q.offN[0] = sx ? 16u : 0u;
    q.offF[0] = sx ? 0u : 16u;
    q.offN[1] = sy ? 48u : 32u;
    q.offF[1] = sy ? 32u : 48u;
    q.offN[2] = sz ? 80u : 64u;
    q.offF[2] = sz ? 64u : 80u;
    q.offK[0] = 4u * uint32_t(kx);
    q.offK[1] = 4u * uint32_t(ky);
    q.offK[2] = 4u * uint32_t(kz);
    q.okx = comp(r.o, kx);
    q.oky = comp(r.o, ky);
    q.okz = comp(r.o, kz);
    int best;
    float bt, bu, bv;
    if (__ballot(!same) == 0)
      packetWalk<COUNT, true>(sc, w, r, q, valid, lane, gt, cap, sRef[wv], sT[wv], best, bt, bu, bv, nN, nT);
    else
      packetWalk<COUNT, false>(sc, w, r, q, valid, lane, gt, cap, sRef[wv], sT[wv], best, bt, bu, bv, nN, nT);
    if (valid) {
      if (best >= 0) {
        const float4* tp = sc.tris + size_t(best) * 3;
        w.hitF[i] = make_float4(bt, bu, bv, 0.f);
        w.hitI[i] = make_int4(__float_as_int(tp[0].w), __float_as_int(tp[1].w), __float_as_int(tp[2].w), best);
      } else {
        w.hitI[i] = make_int4(-1, -1, -1, -1);
      }
    }
   }
  }
  if (COUNT) {
    atomicAdd(&w.cnt64[0], (unsigned long long)nN);
    atomicAdd(&w.cnt64[1], (unsigned long long)nT);
  }
}

// ---------------------------------------------------------------- shadow
// Persistent any-hit loop over the shadow-ray stream: an unoccluded ray
// adds its path's pending NEE radiance.  TREE: 0 binary, 1 4-wide,
// 2 quantized 4-wide.
template <bool COUNT, int TREE, bool XCD>
__device__ __forceinline__ void shadowLoop(const SceneView& sc, const Wavefront& w, Stack& st, int lane,
                                           uint32_t& nNodes, uint32_t& nTris) {
  const uint32_t count = w.ctr[2];
  ChunkFetch f;
  chunkInit<XCD>(f, XCD ? &w.ctr[kXcdCtr + kXcds * 32] : &w.ctr[4], count);
  uint32_t start;
  while (chunkNext<XCD>(f, lane, start)) {
    uint32_t i = start + uint32_t(lane);
    if (i < f.hi) {
      float4 so = w.shO[i], sd = w.shD[i];
      TRay r;
      setupRay(r, xyz(so), xyz(sd));
      int pid = __float_as_int(sd.w);
      bool occluded = TREE ? traceAny4<COUNT, TREE == 2>(sc, r, 0.0f, so.w, st, nNodes, nTris)
                           : traceAny<COUNT>(sc, r, 0.0f, so.w, st, nNodes, nTris);
      if (!occluded) {
        float4 L = w.rad[pid], d = w.ld[i];
        w.rad[pid] = make_float4(L.x + d.x, L.y + d.y, L.z + d.z, L.w);
      }
    }
  }
}

// Any-hit walk in rounds (4-wide float tree), the counterpart of dynRound:
// a lane descends until it holds a leaf (one more is postponed while the
// wave still has lanes at inner nodes), then the leaves are tested.  True
// when the ray is finished; s.hit tells whether it was occluded.
struct AnyState {
  int node, post;
  bool hit;
};
template <bool COUNT>
__device__ __forceinline__ bool anyLeaf(const SceneView& sc, const TRay& r, float tmax, int ref, uint32_t& nT) {
  uint32_t first, cnt;
  decodeLeaf(ref, first, cnt);
  for (uint32_t k = first; k < first + cnt; ++k) {
    const float4* tp = sc.tris + size_t(k) * 3;
    float t, u, v;
    if (COUNT) ++nT;
    if (hitTri(r, tp[0], tp[1], tp[2], 0.0f, tmax, t, u, v)) return true;
  }
  return false;
}
template <bool COUNT>
__device__ __forceinline__ bool anyRound(const SceneView& sc, const TRay& r, float tmax, AnyState& s, Stack& st,
                                         uint32_t& nN, uint32_t& nT) {
  auto next = [&]() { return st.sp > 0 ? st.pop() : kDoneRef; };
  while (s.node >= 0) {
    if (s.node != kEmptyRef) {
      if (COUNT) ++nN;
      Kids k;
      testNode4(sc.nodes4 + size_t(s.node) * 8, r, 0.0f, tmax, k);
      if (k.n > 0) {
        if (k.n > 3) st.push(k.c[3]);
        if (k.n > 2) st.push(k.c[2]);
        if (k.n > 1) st.push(k.c[1]);
        s.node = k.c[0];
      } else {
        s.node = next();
      }
    } else {
      s.node = next();
    }
    if (s.node < 0 && s.node != kDoneRef && s.post == kDoneRef) {
      s.post = s.node;
      s.node = next();
    }
    if (__popcll(__ballot(s.post == kDoneRef)) < kLeafBreak) break;
  }
  if (s.post != kDoneRef) {
    if (anyLeaf<COUNT>(sc, r, tmax, s.post, nT)) {
      s.hit = true;
      return true;
    }
    s.post = kDoneRef;
  }
  if (s.node < 0 && s.node != kDoneRef) {
    if (anyLeaf<COUNT>(sc, r, tmax, s.node, nT)) {
      s.hit = true;
      return true;
    }
    s.node = next();
  }
  return s.node == kDoneRef && s.post == kDoneRef;
}

// Shadow rays with per-lane refill (the fetch of extendDynLoop over the
// shadow stream's queue slices): a lane whose ray has finished takes the
// next one as soon as REFILL lanes of its wave are idle, instead of the
// wave waiting for the slowest of its 64 rays.
template <bool COUNT, int REFILL>
__device__ __forceinline__ void shadowDynLoop(const SceneView& sc, const Wavefront& w, Stack& st, int lane,
                                              uint32_t& nNodes, uint32_t& nTris) {
  ChunkFetch f;
  chunkInit<true>(f, &w.ctr[kXcdCtr + kXcds * 32], w.ctr[2]);
  bool active = false, drained = false;
  int pid = -1;
  uint32_t slot = 0;
  float tmax = 0.0f;
  TRay r;
  AnyState s;
  for (;;) {
    const uint32_t nIdle = uint32_t(__popcll(__ballot(!active)));
    if (!drained && nIdle >= uint32_t(REFILL)) {
      uint32_t i = 0;
      if (fetchIdle(f, !active, lane, drained, i)) {
        const float4 so = w.shO[i], sd = w.shD[i];
        setupRay(r, xyz(so), xyz(sd));
        tmax = so.w;
        pid = __float_as_int(sd.w);
        slot = i;
        s.node = 0;
        s.post = kDoneRef;
        s.hit = false;
        st.sp = 0;
        active = true;
      }
    }
    if (__ballot(active) == 0) {
      if (drained) break;
      continue;
    }
    if (active && anyRound<COUNT>(sc, r, tmax, s, st, nNodes, nTris)) {
      active = false;
      if (!s.hit) {
        float4 L = w.rad[pid], d = w.ld[slot];
        w.rad[pid] = make_float4(L.x + d.x, L.y + d.y, L.z + d.z, L.w);
      }
    }
  }
}

// Two-level closest-hit and shadow walks (rays from the per-XCD queue
// slices, as the flat walks take them), as separate kernels and fused like
// kTraceDyn (a wave that runs out of closest-hit rays takes shadow rays).
//
// The closest-hit walk is resumable, for per-lane refill: its state stays in
// registers between rounds, and each step is one iteration of the walk (a
// TLAS node, the entry into a BLAS with the ray taken to object space, a
// BLAS node or a BLAS leaf, then the pop when the step does not descend).
// It leaves a BLAS when the stack falls back to the depth it had on entry.
// Visit order, culling and tie rule (equal t: smaller global id) are those
// of the flat walk, so the closest hit does not depend on how a lane's steps
// are grouped into rounds.
struct Walk2L {
  int node, cur, base, best, bestGid, bInst;
  float bt, bu, bv;
  uint32_t triBase, blas;
};
constexpr int kSteps2L = 8;    // steps per round between refills
constexpr int kRefill2L = 48;  // idle lanes of a wave that trigger a refill

// One step; true when the walk has finished.
template <bool COUNT>
__device__ __forceinline__ bool step2L(const SceneView& sc, const TRay& rw, TRay& ro, Walk2L& s, Stack& st,
                                       uint32_t& nN, uint32_t& nT) {
  const float tmin = kMinimum, tmax = kInf;
  if (s.node == kDoneRef) {
    // falls through to the pop below
  } else if (s.cur < 0) {
    if (s.node >= 0) {
      if (s.node != kEmptyRef) {
        if (COUNT) ++nN;
        s.node = tlasStep(sc, rw, s.node, tmin, s.best >= 0 ? s.bt : tmax, st);
        if (s.node != kDoneRef) return false;
      }
    } else {
      uint32_t first, cnt;
      decodeLeaf(s.node, first, cnt);
      const uint32_t i = sc.tlasInst[first];
      const DevInstance& in = sc.inst[i];
      toObject(in, rw, ro);
      s.blas = sc.instBlas[i];
      s.cur = int(i);
      s.triBase = in.triBase;
      s.base = st.sp;
      s.node = 0;
      return false;
    }
  } else if (s.node >= 0) {
    if (s.node != kEmptyRef) {
      if (COUNT) ++nN;
      Kids k;
      testNode4(sc.blas[s.blas].nodes4 + size_t(s.node) * 8, ro, tmin, s.best >= 0 ? s.bt : tmax, k);
      if (k.n > 0) {
        if (k.n > 3) st.push(k.c[3], k.t[3]);
        if (k.n > 2) st.push(k.c[2], k.t[2]);
        if (k.n > 1) st.push(k.c[1], k.t[1]);
        s.node = k.c[0];
        return false;
      }
    }
  } else {
    uint32_t first, cnt;
    decodeLeaf(s.node, first, cnt);
    const float4* tris = sc.blas[s.blas].tris;
    for (uint32_t k = first; k < first + cnt; ++k) {
      const float4* tp = tris + size_t(k) * 3;
      const float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
      if (COUNT) ++nT;
      float t, u, v;
      if (hitTri(ro, p0, p1, p2, tmin, s.best >= 0 ? s.bt : tmax, t, u, v)) {
        const int gid = int(s.triBase) + __float_as_int(p2.w);
        if (s.best < 0 || t < s.bt || (t == s.bt && gid < s.bestGid)) {
          s.bt = t, s.bu = u, s.bv = v;
          s.best = int(k);
          s.bestGid = gid;
          s.bInst = s.cur;
        }
      }
    }
  }
  for (;;) {
    if (s.cur >= 0 && st.sp == s.base) s.cur = -1;  // this instance's BLAS is done
    if (st.sp == 0) {
      s.node = kDoneRef;
      return true;
    }
    float tn;
    const int c = st.pop(tn);
    if (s.best < 0 || tn <= s.bt * 1.0000004f) {
      s.node = c;
      return false;
    }
  }
}

// Records the closest hit of queue slot i.  t and barycentrics of the hit
// triangle are evaluated again in world space, on the vertices the flat
// build would store: the object-space test picks the triangle, this gives
// it the flat path's numbers (the object-space ray differs from the world
// ray by rounding, which moves barycentrics by about 1e-4 on small
// triangles).
__device__ __forceinline__ void write2LHit(const SceneView& sc, const Wavefront& w, uint32_t i, const TRay& rw, int inst,
                                           int pos, float t, float u, float v) {
  if (pos < 0) {
    w.hitI[i] = make_int4(-1, -1, -1, -1);
    return;
  }
  const DevBlas& bl = sc.blas[sc.instBlas[inst]];
  const DevInstance& in = sc.inst[inst];
  const float4* tp = bl.tris + size_t(pos) * 3;
  const int prim = __float_as_int(tp[2].w);
  float t2, u2, v2;
  if (hitTri(rw, toWorld(in, tp[0]), toWorld(in, tp[1]), toWorld(in, tp[2]), kMinimum, kInf, t2, u2, v2))
    t = t2, u = u2, v = v2;
  w.hitF[i] = make_float4(t, u, v, 0.f);
  w.hitI[i] = make_int4(inst, int(in.shade), int(in.triBase) + prim, int(bl.shadeBase) + pos);
}

// Persistent closest-hit loop over queue `cur` with per-lane refill (idle
// lanes take new rays from the XCD slices once kRefill2L of them are idle),
// as the flat walk's extendDynLoop; each active lane runs up to kSteps2L
// steps per round.
template <bool COUNT>
__device__ __forceinline__ void extend2LLoop(const SceneView& sc, const Wavefront& w, int cur, Stack& st, int lane,
                                             uint32_t& nNodes, uint32_t& nTris) {
  const uint32_t count = w.ctr[cur];
  const float4 *qo = w.rayO[cur], *qd = w.rayD[cur];
  ChunkFetch f;
  chunkInit<true>(f, &w.ctr[kXcdCtr], count);
  bool active = false, drained = false;
  uint32_t slot = 0;
  TRay rw, ro;
  Walk2L s;
  for (;;) {
    const uint32_t nIdle = uint32_t(__popcll(__ballot(!active)));
    if (!drained && nIdle >= uint32_t(kRefill2L)) {
      uint32_t i = 0;
      if (fetchIdle(f, !active, lane, drained, i)) {
        slot = i;
        setupRay(rw, xyz(qo[i]), xyz(qd[i]));
        s.node = 0;
        s.cur = -1;
        s.base = 0;
        s.best = -1;
        s.bestGid = 0x7fffffff;
        s.bInst = -1;
        s.bt = kInf;
        s.bu = s.bv = 0.f;
        s.triBase = 0;
        s.blas = 0;
        st.sp = 0;
        active = true;
      }
    }
    if (__ballot(active) == 0) {
      if (drained) break;
      continue;
    }
    if (active) {
      bool done = false;
      for (int k = 0; k < kSteps2L && !done; ++k) done = step2L<COUNT>(sc, rw, ro, s, st, nNodes, nTris);
      if (done) {
        active = false;
        write2LHit(sc, w, slot, rw, s.bInst, s.best, s.bt, s.bu, s.bv);
      }
    }
  }
}
// The any-hit walk in the same resumable form; step result 0: walking,
// 1: finished without a hit, 2: occluded.
struct Any2L {
  int node, cur, base;
  uint32_t blas;
};
template <bool COUNT>
__device__ __forceinline__ int stepAny2L(const SceneView& sc, const TRay& rw, TRay& ro, float tmax, Any2L& s,
                                         Stack& st, uint32_t& nN, uint32_t& nT) {
  const float tmin = 0.0f;
  if (s.node == kDoneRef) {
    // falls through to the pop below
  } else if (s.cur < 0) {
    if (s.node >= 0) {
      if (s.node != kEmptyRef) {
        if (COUNT) ++nN;
        s.node = tlasStep(sc, rw, s.node, tmin, tmax, st);
        if (s.node != kDoneRef) return 0;
      }
    } else {
      uint32_t first, cnt;
      decodeLeaf(s.node, first, cnt);
      const uint32_t i = sc.tlasInst[first];
      toObject(sc.inst[i], rw, ro);
      s.blas = sc.instBlas[i];
      s.cur = int(i);
      s.base = st.sp;
      s.node = 0;
      return 0;
    }
  } else if (s.node >= 0) {
    if (s.node != kEmptyRef) {
      if (COUNT) ++nN;
      Kids k;
      testNode4(sc.blas[s.blas].nodes4 + size_t(s.node) * 8, ro, tmin, tmax, k);
      if (k.n > 0) {
        if (k.n > 3) st.push(k.c[3], k.t[3]);
        if (k.n > 2) st.push(k.c[2], k.t[2]);
        if (k.n > 1) st.push(k.c[1], k.t[1]);
        s.node = k.c[0];
        return 0;
      }
    }
  } else {
    uint32_t first, cnt;
    decodeLeaf(s.node, first, cnt);
    const float4* tris = sc.blas[s.blas].tris;
    for (uint32_t k = first; k < first + cnt; ++k) {
      const float4* tp = tris + size_t(k) * 3;
      float t, u, v;
      if (COUNT) ++nT;
      if (hitTri(ro, tp[0], tp[1], tp[2], tmin, tmax, t, u, v)) return 2;
    }
  }
  if (s.cur >= 0 && st.sp == s.base) s.cur = -1;  // this instance's BLAS is done
  if (st.sp == 0) return 1;
  float tn;
  s.node = st.pop(tn);
  return 0;
}

// Persistent any-hit loop over the shadow-ray queue with per-lane refill,
// as extend2LLoop: an unoccluded ray adds its path's pending NEE radiance.
template <bool COUNT>
__device__ __forceinline__ void shadow2LLoop(const SceneView& sc, const Wavefront& w, Stack& st, int lane,
                                             uint32_t& nNodes, uint32_t& nTris) {
  const uint32_t count = w.ctr[2];
  ChunkFetch f;
  chunkInit<true>(f, &w.ctr[kXcdCtr + kXcds * 32], count);
  bool active = false, drained = false;
  uint32_t slot = 0;
  int pid = 0;
  float tmax = 0.f;
  TRay rw, ro;
  Any2L s;
  for (;;) {
    const uint32_t nIdle = uint32_t(__popcll(__ballot(!active)));
    if (!drained && nIdle >= uint32_t(kRefill2L)) {
      uint32_t i = 0;
      if (fetchIdle(f, !active, lane, drained, i)) {
        slot = i;
        const float4 so = w.shO[i], sd = w.shD[i];
        setupRay(rw, xyz(so), xyz(sd));
        tmax = so.w;
        pid = __float_as_int(sd.w);
        s.node = 0;
        s.cur = -1;
        s.base = 0;
        s.blas = 0;
        st.sp = 0;
        active = true;
      }
    }
    if (__ballot(active) == 0) {
      if (drained) break;
      continue;
    }
    if (active) {
      int r = 0;
      for (int k = 0; k < kSteps2L && r == 0; ++k) r = stepAny2L<COUNT>(sc, rw, ro, tmax, s, st, nNodes, nTris);
      if (r != 0) {
        active = false;
        if (r == 1) {
          const float4 L = w.rad[pid], d = w.ld[slot];
          w.rad[pid] = make_float4(L.x + d.x, L.y + d.y, L.z + d.z, L.w);
        }
      }
    }
  }
}
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void kExtend2L(SceneView sc, Wavefront w, int cur) {
  __shared__ int lds[kStackLds * kTraceBlock];
  __shared__ float ldsT[kStackLds * kTraceBlock];
  Stack st;
  initStack<true, kStackLds>(st, lds, ldsT, w);
  uint32_t nNodes = 0, nTris = 0;
  extend2LLoop<COUNT>(sc, w, cur, st, __lane_id(), nNodes, nTris);
  if (COUNT) {
    atomicAdd(&w.cnt64[0], (unsigned long long)nNodes);
    atomicAdd(&w.cnt64[1], (unsigned long long)nTris);
  }
}
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void kShadow2L(SceneView sc, Wavefront w) {
  __shared__ int lds[kStackLds * kTraceBlock];
  __shared__ float ldsT[kStackLds * kTraceBlock];
  Stack st;
  initStack<true, kStackLds>(st, lds, ldsT, w);
  uint32_t nNodes = 0, nTris = 0;
  shadow2LLoop<COUNT>(sc, w, st, __lane_id(), nNodes, nTris);
  if (COUNT) {
    atomicAdd(&w.cnt64[2], (unsigned long long)nNodes);
    atomicAdd(&w.cnt64[3], (unsigned long long)nTris);
  }
}
// COUNT: as kTraceDyn (totals and the fused launches' own counters).
template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void kTrace2L(SceneView sc, Wavefront w, int cur) {
  __shared__ int lds[kStackLds * kTraceBlock];
  __shared__ float ldsT[kStackLds * kTraceBlock];
  if (COUNT && blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd(&w.cnt64[11], (unsigned long long)w.ctr[cur]);
    atomicAdd(&w.cnt64[12], (unsigned long long)w.ctr[2]);
  }
  Stack st;
  initStack<true, kStackLds>(st, lds, ldsT, w);
  uint32_t eN = 0, eT = 0, sN = 0, sT = 0;
  extend2LLoop<COUNT>(sc, w, cur, st, __lane_id(), eN, eT);
  shadow2LLoop<COUNT>(sc, w, st, __lane_id(), sN, sT);
  if (COUNT) {
    const unsigned long long v[4] = {eN, eT, sN, sT};
    for (int k = 0; k < 4; ++k) {
      atomicAdd(&w.cnt64[k], v[k]);
      atomicAdd(&w.cnt64[7 + k], v[k]);
    }
  }
}

template <bool COUNT, int TREE, bool XCD>
__global__ __launch_bounds__(kTraceBlock) void kShadow(SceneView sc, Wavefront w) {
  __shared__ int lds[kStackLds * kTraceBlock];
  Stack st;
  initStack<false>(st, lds, nullptr, w);
  uint32_t nNodes = 0, nTris = 0;
  if (TREE == 1 && XCD) shadowDynLoop<COUNT, kShadowRefill>(sc, w, st, __lane_id(), nNodes, nTris);
  else shadowLoop<COUNT, TREE, XCD>(sc, w, st, __lane_id(), nNodes, nTris);
  if (COUNT) {
    atomicAdd(&w.cnt64[2], (unsigned long long)nNodes);
    atomicAdd(&w.cnt64[3], (unsigned long long)nTris);
  }
}

// The shadow rays of one bounce and the extension rays of the next in one
// persistent launch (they touch disjoint path state: the shadow pass adds
// to rad, the extension pass writes the hit records).  A wave that runs out
// of extension rays goes on to shadow rays while the others finish, so the
// latency tail of each pass (its slowest rays, when few waves still have
// work) overlaps the other pass instead of idling the device twice.
// COUNT: traversal work is added to the totals (cnt64[0..3]) and to the
// fused launches' own counters (cnt64[7..12]: extension nodes, triangles,
// shadow nodes, triangles, extension rays, shadow rays).
template <bool COUNT, int NL, int REFILL>
__global__ __launch_bounds__(kTraceBlock) void kTraceDyn(SceneView sc, Wavefront w, int cur) {
  __shared__ int lds[NL * kTraceBlock];
  __shared__ float ldsT[NL * kTraceBlock];
  const int lane = __lane_id();
  uint32_t eN = 0, eT = 0, sN = 0, sT = 0;
  if (COUNT && blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd(&w.cnt64[11], (unsigned long long)w.ctr[cur]);
    atomicAdd(&w.cnt64[12], (unsigned long long)w.ctr[2]);
  }
  Stack st;
  initStack<true, NL>(st, lds, ldsT, w);
  extendDynLoop<COUNT, REFILL>(sc, w, cur, st, lane, eN, eT);
  Stack ss;
  initStack<false, NL>(ss, lds, nullptr, w);
  shadowDynLoop<COUNT, kShadowRefill>(sc, w, ss, lane, sN, sT);
  if (COUNT) {
    const unsigned long long v[4] = {eN, eT, sN, sT};
    for (int k = 0; k < 4; ++k) {
      atomicAdd(&w.cnt64[k], v[k]);
      atomicAdd(&w.cnt64[7 + k], v[k]);
    }
  }
}

// ------------------------------------------- 4-wide walk with ray refill
// Persistent waves keep every lane busy: a lane whose ray has finished
// takes a new one from the queue as soon as kRefill lanes of its wave are
// idle (one atomic per refill), and each loop iteration advances every busy
// lane by one step -- one triangle test or one node visit -- so lanes at
// leaves and lanes at inner nodes stay in lockstep instead of the wave
// waiting for its slowest ray.
constexpr int kRefill = 16;

struct Trav {
  TRay r;
  int node;  // next inner node to visit, kEmptyRef = none
  uint32_t leafFirst, leafLeft;
  int best, bestGid;
  float bt, bu, bv, tmax;
};

__device__ __forceinline__ void takeRef(Trav& s, int ref) {
  if (ref < 0) {
    decodeLeaf(ref, s.leafFirst, s.leafLeft);
    s.node = kEmptyRef;
  } else {
    s.node = ref;
  }
}

__device__ __forceinline__ void startRay(Trav& s, Stack& st, f3 o, f3 d, float tmax) {
  setupRay(s.r, o, d);
  s.node = 0;
  s.leafFirst = s.leafLeft = 0;
  s.best = -1;
  s.bestGid = 0x7fffffff;
  s.bt = s.tmax = tmax;
  s.bu = s.bv = 0.f;
  st.sp = 0;
}

// One step of the closest-hit walk; true when the ray has finished.  Same
// visit order, culling and tie rule as traceClosest4.
template <bool COUNT>
__device__ __forceinline__ bool stepClosest4(const SceneView& sc, Trav& s, Stack& st, uint32_t& nN, uint32_t& nT) {
  if (s.leafLeft) {
    const float4* tp = sc.tris + size_t(s.leafFirst) * 3;
    float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
    if (COUNT) ++nT;
    float t, u, v;
    if (hitTri(s.r, p0, p1, p2, kMinimum, s.best >= 0 ? s.bt : s.tmax, t, u, v)) {
      int gid = __float_as_int(p2.w);
      if (s.best < 0 || t < s.bt || (t == s.bt && gid < s.bestGid)) {
        s.bt = t;
        s.bu = u;
        s.bv = v;
        s.best = int(s.leafFirst);
        s.bestGid = gid;
      }
    }
    ++s.leafFirst;
    --s.leafLeft;
  } else if (s.node != kEmptyRef) {
    if (COUNT) ++nN;
    Kids k;
    testNode4(sc.nodes4 + size_t(s.node) * 8, s.r, kMinimum, s.best >= 0 ? s.bt : s.tmax, k);
    s.node = kEmptyRef;
    if (k.n > 0) {
      if (k.n > 3) st.push(k.c[3], k.t[3]);
      if (k.n > 2) st.push(k.c[2], k.t[2]);
      if (k.n > 1) st.push(k.c[1], k.t[1]);
      takeRef(s, k.c[0]);
    }
  }
  if (s.leafLeft == 0 && s.node == kEmptyRef) {
    while (st.sp > 0) {
      float tn;
      int c = st.pop(tn);
      if (s.best < 0 || tn <= s.bt * 1.0000004f) {
        takeRef(s, c);
        if (s.leafLeft || s.node != kEmptyRef) break;
      }
    }
    return s.leafLeft == 0 && s.node == kEmptyRef;
  }
  return false;
}

// One step of the any-hit walk: 0 running, 1 finished unoccluded, 2 occluded.
template <bool COUNT>
__device__ __forceinline__ int stepAny4(const SceneView& sc, Trav& s, Stack& st, uint32_t& nN, uint32_t& nT) {
  if (s.leafLeft) {
    const float4* tp = sc.tris + size_t(s.leafFirst) * 3;
    if (COUNT) ++nT;
    float t, u, v;
    if (hitTri(s.r, tp[0], tp[1], tp[2], 0.0f, s.tmax, t, u, v)) return 2;
    ++s.leafFirst;
    --s.leafLeft;
  } else if (s.node != kEmptyRef) {
    if (COUNT) ++nN;
    Kids k;
    testNode4(sc.nodes4 + size_t(s.node) * 8, s.r, 0.0f, s.tmax, k);
    s.node = kEmptyRef;
    if (k.n > 0) {
      if (k.n > 3) st.push(k.c[3]);
      if (k.n > 2) st.push(k.c[2]);
      if (k.n > 1) st.push(k.c[1]);
      takeRef(s, k.c[0]);
    }
  }
  if (s.leafLeft == 0 && s.node == kEmptyRef) {
    while (st.sp > 0) {
      takeRef(s, st.pop());
      if (s.leafLeft || s.node != kEmptyRef) break;
    }
    if (s.leafLeft == 0 && s.node == kEmptyRef) return 1;
  }
  return 0;
}

// Hands queue entries to the idle lanes of the wave when at least kRefill
// of them are idle; `drained` is set once the queue has been handed out.
__device__ __forceinline__ uint32_t refill(uint32_t* fetchCtr, uint32_t count, bool active, bool& drained,
                                           bool& got) {
  got = false;
  uint64_t idle = __ballot(!active);
  int nIdle = __popcll(idle);
  if (drained || nIdle < kRefill) return 0;
  int lane = __lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(fetchCtr, uint32_t(nIdle));
  base = __shfl(base, 0);
  if (base + uint32_t(nIdle) >= count) drained = true;
  uint32_t i = base + uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
  got = !active && i < count;
  return i;
}

template <bool COUNT, int NL>
__global__ __launch_bounds__(kTraceBlock) void kExtendWide(SceneView sc, Wavefront w, int cur) {
  __shared__ int lds[NL * kTraceBlock];
  __shared__ float ldsT[NL * kTraceBlock];
  Stack st;
  initStack<true, NL>(st, lds, ldsT, w);
  const uint32_t count = w.ctr[cur];
  const float4 *qo = w.rayO[cur], *qd = w.rayD[cur];
  uint32_t nNodes = 0, nTris = 0;
  Trav s;
  uint32_t slot = 0;
  bool active = false, drained = false;
  for (;;) {
    bool got;
    uint32_t i = refill(&w.ctr[3], count, active, drained, got);
    if (got) {
      slot = i;
      startRay(s, st, xyz(qo[i]), xyz(qd[i]), kInf);
      active = true;
    }
    if (__ballot(active) == 0) {
      if (drained) break;
      continue;
    }
    if (active && stepClosest4<COUNT>(sc, s, st, nNodes, nTris)) {
      active = false;
      if (s.best >= 0) {
        const float4* tp = sc.tris + size_t(s.best) * 3;
        w.hitF[slot] = make_float4(s.bt, s.bu, s.bv, 0.f);
        w.hitI[slot] = make_int4(__float_as_int(tp[0].w), __float_as_int(tp[1].w), __float_as_int(tp[2].w), s.best);
      } else {
        w.hitI[slot] = make_int4(-1, -1, -1, -1);
      }
    }
  }
  if (COUNT) {
    atomicAdd(&w.cnt64[0], (unsigned long long)nNodes);
    atomicAdd(&w.cnt64[1], (unsigned long long)nTris);
  }
}

template <bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void kShadowWide(SceneView sc, Wavefront w) {
  __shared__ int lds[kStackLds * kTraceBlock];
  Stack st;
  initStack<false>(st, lds, nullptr, w);
  const uint32_t count = w.ctr[2];
  uint32_t nNodes = 0, nTris = 0;
  Trav s;
  int pid = -1;
  uint32_t slot = 0;
  bool active = false, drained = false;
  for (;;) {
    bool got;
    uint32_t i = refill(&w.ctr[4], count, active, drained, got);
    if (got) {
      float4 so = w.shO[i], sd = w.shD[i];
      pid = __float_as_int(sd.w);
      slot = i;
      startRay(s, st, xyz(so), xyz(sd), so.w);
      active = true;
    }
    if (__ballot(active) == 0) {
      if (drained) break;
      continue;
    }
    if (active) {
      int r = stepAny4<COUNT>(sc, s, st, nNodes, nTris);
      if (r) {
        active = false;
        if (r == 1) {
          float4 L = w.rad[pid], d = w.ld[slot];
          w.rad[pid] = make_float4(L.x + d.x, L.y + d.y, L.z + d.z, L.w);
        }
      }
    }
  }
  if (COUNT) {
    atomicAdd(&w.cnt64[2], (unsigned long long)nNodes);
    atomicAdd(&w.cnt64[3], (unsigned long long)nTris);
  }
}

// ----------------------------------------------------------------- shade
struct Hit {
  int lightId;
  int inst;
  float uvx, uvy;
  f3 pos, V, N, ffN, X, Y;
  uint32_t mat;
};

// getHitState + configureShadingFrame, rchit_layouts.glsl:61-95.
// `leaf`: the triangle's position in leaf order, whose shading record holds
// the same vertex data as the mesh's index triple and vertices.
// `shade`: the instance's shade word (emitter or material id), carried by
// the hit so that the material can be fetched without the instance record.
__device__ __forceinline__ void hitState(const SceneView& sc, int inst, uint32_t shade, int leaf, float bu, float bv,
                                         f3 rayDir, uint32_t useFaceNormal, Hit& h) {
  const DevInstance& in = sc.inst[inst];
  const float4* sr = sc.shadeTris + size_t(leaf) * 6;
  float4 a0 = sr[0], b0 = sr[1], a1 = sr[2], b1 = sr[3], a2 = sr[4], b2 = sr[5];
  f3 ba = f3{1.0f - bu - bv, bu, bv};
  h.lightId = (shade & kShadeLight) ? int(shade & kShadeIdMask) : -1;
  h.uvx = a0.w * ba.x + a1.w * ba.y + a2.w * ba.z;
  h.uvy = b0.w * ba.x + b1.w * ba.y + b2.w * ba.z;
  f3 op = xyz(a0) * ba.x + xyz(a1) * ba.y + xyz(a2) * ba.z;
  const float* m = in.o2w;
  h.pos = f3{m[0] * op.x + m[1] * op.y + m[2] * op.z + m[3], m[4] * op.x + m[5] * op.y + m[6] * op.z + m[7],
             m[8] * op.x + m[9] * op.y + m[10] * op.z + m[11]};
  f3 n = xyz(b0) * ba.x + xyz(b1) * ba.y + xyz(b2) * ba.z;
  const float* q = in.w2o;
  f3 nw = f3{n.x * q[0] + n.y * q[4] + n.z * q[8], n.x * q[1] + n.y * q[5] + n.z * q[9],
             n.x * q[2] + n.y * q[6] + n.z * q[10]};
  h.N = makeNormal(nw);
  h.V = makeNormal(-rayDir);
  h.mat = shade & kShadeIdMask;
  h.inst = inst;
  if (useFaceNormal == 1) h.N = mk1(0.f);  // geoN is never assigned (rchit_layouts.glsl:62)
  basis(h.N, h.X, h.Y);
  h.ffN = dot(h.N, h.V) > 0.0f ? h.N : -h.N;
}

__device__ __forceinline__ void normalMap(const SceneView& sc, Hit& h, int texId, uint32_t useFaceNormal) {
  float4 c = tex2D(sc.textures[texId], h.uvx, h.uvy);
  f3 n = 2.0f * xyz(c) - mk1(1.f);
  h.N = toWorld(h.X, h.Y, h.N, n);
  h.N = makeNormal(h.N);
  if (useFaceNormal == 1) h.N = mk1(0.f);
  basis(h.N, h.X, h.Y);
  h.ffN = dot(h.N, h.V) > 0.0f ? h.N : -h.N;
}

struct LSample {
  f3 d, n;
  float dist, pdf;
  uint32_t flags;
};

// sampleEnvmap, sample_light.glsl:106-129
__device__ __forceinline__ f3 sampleEnv(const SceneView& sc, float r1, float r2, const float* envT, float resx,
                                        float resy, float intensity, f3& L, float& pdf) {
  uint32_t W = sc.env.w, H = sc.env.h;
  float uvy = texMarginal(sc.marg, W, H, r1).x;
  float uvx = texCondX(sc.condX, W, H, r2, uvy);
  const float4 e = texEnv(sc.env, uvx, uvy);  // radiance and conditional pdf of the direction
  pdf = e.w * texMarginal(sc.marg, W, H, uvy).y;
  float phi = uvx * kTwoPi;
  float theta = uvy * kPi;
  if (sinf(theta) == 0.0f) pdf = 0.0f;
  pdf = (pdf * resx * resy) / (kTwoPi * kPi * sinf(theta));
  L = f3{-sinf(theta) * cosf(phi), cosf(theta), -sinf(theta) * sinf(phi)};
  L = transformDirection(envT, L);
  return intensity * xyz(e);
}
__device__ __forceinline__ void envUV(const float* envT, f3 L, float& u, float& v, float& theta) {
  L = makeNormal(transformVectorT(envT, L));
  theta = acosf(clampf(L.y, -1.0f, 1.0f));
  u = (kPi + atan2f(L.z, L.x)) * kInv2Pi;
  v = theta * kInvPi;
}
// evalEnvmap, sample_light.glsl:97-104, and pdfEnvmap, sample_light.glsl:84-95,
// of one direction from one texel lookup (the miss shader needs both).
__device__ __forceinline__ f3 evalEnvPdf(const SceneView& sc, const float* envT, float intensity, float resx,
                                         float resy, f3 L, float& pdfOut) {
  float u, v, theta;
  envUV(envT, L, u, v, theta);
  const float4 e = texEnv(sc.env, u, v);
  float pdf = e.w * texMarginal(sc.marg, sc.env.w, sc.env.h, v).y;
  float s = sinf(theta);
  pdfOut = s == 0.0f ? 0.0f : (pdf * resx * resy) / (kTwoPi * kPi * s);
  return intensity * xyz(e);
}

// Sun & sky, utils/sun_and_sky.glsl:20-533 (NVIDIA's MDL sun & sky:
// Preetham/Perez sky, attenuated sun disk with glow, ground below the
// horizon).  The file's own M_PI is 3.1415926535f.
constexpr float kSkyPi = 3.1415926535f;
__device__ __forceinline__ float skySmoothstep(float e0, float e1, float x) {
  float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ f3 skyTweak(f3 d, int yUp, float hh) {  // arch_vectortweak
  f3 o = yUp == 1 ? f3{d.x, d.z, d.y} : d;
  if (hh != 0.0f) {
    o.z -= hh;
    o = makeNormal(o);
  }
  return o;
}
__device__ __forceinline__ float skyPerez(float A, float B, float C, float D, float E, float ct, float gamma,
                                          float cg, float ts, float cts) {
  return ((1.0f + A * expf(B / ct)) * (1.0f + C * expf(D * gamma) + E * cg * cg)) /
         ((1.0f + A * expf(B / 1.0f)) * (1.0f + C * expf(D * ts) + E * cts * cts));
}
// calc_env_color with sky_luminance and sky_color_xyz, sun_and_sky.glsl:134-244
__device__ f3 skyEnvColor(f3 sun, f3 dir, float T) {
  const float cts = sun.z, ts = acosf(cts), ct = dir.z;
  float chi = (4.0f / 9.0f - T / 120.0f) * (kSkyPi - 2.0f * ts);
  float lum = 1000.0f * ((4.0453f * T - 4.9710f) * tanf(chi) - 0.2155f * T + 2.4192f);
  float cg = dot(sun, dir);
  float cgL = cg < 0.0f ? 0.0f : cg;  // sky_luminance clamps below at 0, sky_color_xyz does not
  if (cgL > 1.0f) cgL = 2.0f - cgL;
  if (cg > 1.0f) cg = 2.0f - cg;
  lum *= skyPerez(0.178721f * T - 1.463037f, -0.355402f * T + 0.427494f, -0.022669f * T + 5.325056f,
                  0.120647f * T - 2.577052f, -0.066967f * T + 0.370275f, ct, acosf(cgL), cgL, ts, cts);
  float gamma = acosf(cg);
  float t2 = T * T, ts2 = ts * ts, ts3 = ts2 * ts;
  float zx = (0.001650f * ts3 - 0.003742f * ts2 + 0.002088f * ts + 0.0f) * t2 +
             (-0.029028f * ts3 + 0.063773f * ts2 - 0.032020f * ts + 0.003948f) * T +
             (0.116936f * ts3 - 0.211960f * ts2 + 0.060523f * ts + 0.258852f);
  float zy = (0.002759f * ts3 - 0.006105f * ts2 + 0.003162f * ts + 0.0f) * t2 +
             (-0.042149f * ts3 + 0.089701f * ts2 - 0.041536f * ts + 0.005158f) * T +
             (0.153467f * ts3 - 0.267568f * ts2 + 0.066698f * ts + 0.266881f);
  float x = skyPerez(-0.019257f * T - (0.29f - powf(cts, 0.5f) * 0.09f), -0.066513f * T + 0.000818f,
                     -0.000417f * T + 0.212479f, -0.064097f * T - 0.898875f, -0.003251f * T + 0.045178f, ct, gamma,
                     cg, ts, cts);
  float y = skyPerez(-0.016698f * T - 0.260787f, -0.094958f * T + 0.009213f, -0.007928f * T + 0.210230f,
                     -0.044050f * T - 1.653694f, -0.010922f * T + 0.052919f, ct, gamma, cg, ts, cts);
  x = zx * ((x * 1.0f) + (1.0f - 1.0f));
  y = zy * ((y * 1.0f) + (1.0f - 1.0f));
  float X = (x / y) * lum, Y = lum, Z = ((1.0f - x - y) / y) * lum;
  f3 c = f3{3.241f * X - 1.537f * Y - 0.499f * Z, -0.969f * X + 1.876f * Y + 0.042f * Z,
            0.056f * X - 0.204f * Y + 1.057f * Z};
  return c * kSkyPi;
}
// calc_sun_color, sun_and_sky.glsl:108-131
__device__ f3 skySunColor(f3 sd, float T) {
  if (!(sd.z > 0.0f)) return mk1(0.f);
  float m = 1.0f / (sd.z + 0.15f * powf(93.885f - acosf(sd.z) * 180.0f / kSkyPi, -1.253f));
  float beta = 0.04608f * T - 0.04586f;
  const float ko[3] = {12.0f, 8.5f, 0.9f}, wl[3] = {0.610f, 0.550f, 0.470f};
  const float sol[3] = {1.0f * 127500.0f / 0.9878f, 0.992f * 127500.0f / 0.9878f, 0.911f * 127500.0f / 0.9878f};
  float o[3];
  for (int k = 0; k < 3; ++k) {
    float ta = expf(-m * beta * powf(wl[k], -1.3f));
    float to = expf(-m * ko[k] * 0.0035f);
    float tr = expf(-m * 0.008735f * powf(wl[k], -4.08f));
    o[k] = tr * ta * to * sol[k];
  }
  return f3{o[0], o[1], o[2]};
}

// The direction-independent part of sun_and_sky (one thread per record).
__global__ void kSkyPrepare(DevSky* k) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const GpuSunAndSky ss = k->ss;
  k->hh = ss.horizon_height / 10.0f;
  float haze = 2.0f + ss.haze;
  if (haze < 2.0f) haze = 2.0f;
  k->haze = haze;
  {  // tweak_saturation, sun_and_sky.glsl:264-277
    float sat = ss.saturation, low = powf(sat, 3.0f), out = 1.0f;
    if (sat <= 1.0f) {
      float h = clampf((haze - 2.0f) / 15.0f, 0.0f, 1.0f);
      h = powf(h, 3.0f);
      out = sat * (1.0f - h) + low * h;
    }
    k->sat = out;
  }
  f3 rs = ld3(ss.rgb_unit_conversion);
  if (luminance(rs) < 0.0f) rs = mk1(1.0f / 80000.0f);
  rs = rs * ss.multiplier;
  f3 sun = skyTweak(makeNormal(ld3(ss.sun_direction)), ss.y_is_up, k->hh);
  f3 realSun = sun;
  float factor = 1.0f;
  if (sun.z < 0.001f) {
    if (sun.z < 0.0f) {  // night_brightness_adjustment, sun_and_sky.glsl:398-405
      const float lmt = 0.30901699437494742410229341718282f;
      if (sun.z <= -lmt) {
        factor = 0.0f;
      } else {
        float f = (sun.z + lmt) / lmt;
        f *= f;
        f *= f;
        factor = f;
      }
    }
    sun.z = 0.001f;
    sun = makeNormal(sun);
  }
  f3 up = skySunColor(sun, haze), dn = skySunColor(sun, 2.0f);
  // calc_irrad, sun_and_sky.glsl:246-262: 5 x 5 stratified cosine directions
  // (mi_reflection_dir_diffuse_x about +z), float loop counters.
  f3 irr = mk1(0.f);
  for (float u = 1.0f / 10.0f; u < 1.0f; u += 1.0f / 5.0f)
    for (float v = 1.0f / 10.0f; v < 1.0f; v += 1.0f / 5.0f) {
      float lx = 2.0f * u - 1.0f, ly = 2.0f * v - 1.0f, r, phi;
      const float q = kSkyPi / 4.0f;
      if (lx == 0.0f && ly == 0.0f) r = 0.0f, phi = 0.0f;
      else if (lx > -ly) {
        if (lx > ly) r = lx, phi = q * (1.0f + ly / lx);
        else r = ly, phi = q * (3.0f - lx / ly);
      } else {
        if (lx < ly) r = -lx, phi = q * (5.0f + ly / lx);
        else r = -ly, phi = q * (7.0f - lx / ly);
      }
      float x = r * cosf(phi), y = r * sinf(phi);
      float z2 = 1.0f - x * x - y * y;
      float z = z2 > 0.0f ? sqrtf(z2) : 0.0f;
      // xyz2dir about n = +z: u = (1, 0, 0), v = n x u = (0, 1, 0)
      f3 d = x * f3{1.0f, 0.0f, 0.0f} + y * f3{0.0f, 1.0f, 0.0f} + z * f3{0.0f, 0.0f, 1.0f};
      irr = irr + skyEnvColor(sun, d, 2.0f);
    }
  irr = irr / 25.0f;
  f3 ground = ld3(ss.ground_color) * (irr + dn * sun.z) * rs;
  if (factor < 1.0f) ground = ground * factor;
  float sd = 1.0f, sg = 1.0f;
  if (ss.physically_scaled_sun == 1) {  // calc_physical_scale, sun_and_sky.glsl:315-396
    float diskR = 0.00465f * ss.sun_disk_scale, glowR = diskR * 10.0f;
    float glowInt = ss.sun_glow_intensity * ((4.0f * kSkyPi) - (24.0f * kSkyPi) / (glowR * glowR) +
                                             (24.0f * kSkyPi) * sinf(glowR) / (glowR * glowR * glowR));
    float target = ss.sun_disk_intensity * kSkyPi;
    float maxGlow = 0.5f * target;
    if (glowInt > maxGlow) {
      sg *= maxGlow / glowInt;
      target -= maxGlow;
    } else {
      target -= glowInt;
    }
    float area = 2.0f * kSkyPi * (1.0f - cosf(diskR));
    float targetI = target / area;
    float actual = ss.sun_disk_intensity * 100.0f * (1.0f * area) / area;
    sd = targetI == 0.0f ? 0.0f : targetI / actual;
  }
  for (int i = 0; i < 3; ++i) {
    k->sun[i] = (&sun.x)[i];
    k->realSun[i] = (&realSun.x)[i];
    k->rgbScale[i] = (&rs.x)[i];
    k->sunUp[i] = (&up.x)[i];
    k->sunDown[i] = (&dn.x)[i];
    k->ground[i] = (&ground.x)[i];
  }
  k->factor = factor;
  k->diskScale = sd;
  k->glowScale = sg;
  k->inUse = ss.in_use;
}

// sun_and_sky, sun_and_sky.glsl:407-531, per direction.
__device__ f3 skyEval(const DevSky& k, f3 inDir) {
  const GpuSunAndSky& ss = k.ss;
  if (ss.multiplier <= 0.0f) return mk1(0.f);
  f3 dir = skyTweak(inDir, ss.y_is_up, k.hh);
  const float down = dir.z;
  const f3 realDir = dir;
  if (dir.z < 0.001f) {
    dir.z = 0.001f;
    dir = makeNormal(dir);
  }
  f3 tint = mk1(0.f);
  if (k.factor > 0.0f) {
    tint = skyEnvColor(ld3(k.sun), dir, k.haze);
    if (k.factor < 1.0f) tint = tint * k.factor;
  }
  if (ss.sun_disk_intensity > 0.0f && ss.sun_disk_scale > 0.0f) {
    float angle = acosf(dot(realDir, ld3(k.realSun)));
    float radius = 0.00465f * ss.sun_disk_scale * 10.0f;
    if (angle < radius) {
      float f = (1.0f - angle / radius) * 10.0f;
      f = powf(f / 10.0f, 3.0f) * 2.0f * ss.sun_glow_intensity * k.glowScale +
          skySmoothstep(8.5f, 9.5f + (k.haze / 50.0f), f) * 100.0f * ss.sun_disk_intensity * k.diskScale;
      tint = tint + (down > 0.0f ? ld3(k.sunUp) : ld3(k.sunDown)) * f;
    }
  }
  f3 out = tint * ld3(k.rgbScale);
  float nightF = 1.0f;
  if (down <= 0.0f) {
    f3 dc = ld3(k.ground);
    float blur = ss.horizon_blur / 10.0f;
    if (blur > 0.0f) {
      float dn = -down / blur;
      if (dn > 1.0f) dn = 1.0f;
      dn = skySmoothstep(0.0f, 1.0f, dn);
      out = out * (1.0f - dn) + dc * dn;
      nightF = 1.0f - dn;
    } else {
      out = dc;
      nightF = 0.0f;
    }
  }
  {  // arch_colortweak, sun_and_sky.glsl:291-313
    float I = luminance(out), sat = k.sat, red = ss.redblueshift;
    out = sat <= 0.0f ? mk1(I) : out * sat + mk1(I * (1.0f - sat));
    out = out * f3{1.0f + red, 1.0f, 1.0f - red};
  }
  if (nightF > 0.0f) {
    f3 night = ld3(ss.night_color) * nightF;
    out = f3{fmaxf(out.x, night.x), fmaxf(out.y, night.y), fmaxf(out.z, night.z)};
  }
  return out * kSkyPi;
}

// sampleRectLight / sampleTriangleLight / samplePointLight /
// sampleDistantLight, sample_light.glsl:10-82
__device__ __forceinline__ f3 sampleOneLight(float r0, float r1, const GpuLight& L, f3 p, LSample& s) {
  int type = L.type;
  f3 lu = ld3(L.u), lv = ld3(L.v);
  if (type == LightTypeRect || type == LightTypeTriangle) {
    float b = type == LightTypeRect ? r1 : (1.0f - r0) * r1;
    f3 lp = ld3(L.position) + lu * r0 + lv * b;
    s.d = lp - p;
    s.dist = length(s.d);
    float distSq = s.dist * s.dist;
    s.d = s.d / s.dist;
    s.n = makeNormal(cross(lu, lv));
    s.pdf = distSq / (L.area * fabsf(dot(s.n, s.d)) + kEps);
    s.flags = kLightArea;
    return ld3(L.radiance);
  }
  if (type == LightTypeDirectional) {
    s.d = makeNormal(ld3(L.direction));
    s.n = -s.d;
    s.dist = kInf;
    s.pdf = 1.0f;
    s.flags = kLightDelta;
    return ld3(L.radiance);
  }
  if (type == LightTypePoint) {
    s.d = ld3(L.position) - p;
    s.n = -s.d;
    s.dist = length(s.d);
    float distSq = s.dist * s.dist;
    s.d = s.d / (s.dist + kEps);
    s.pdf = 1.0f;
    s.flags = kLightDelta;
    return ld3(L.radiance) / (distSq + kEps);
  }
  return mk1(0.f);
}

struct PathCtx {
  const SceneView* sc;
  const ShotParams* sp;
  uint32_t seed;
  f3 ro, rd, T, L;
  float bpdf;
  uint32_t bflags, depth;
  bool stop;
  // NEE result
  bool shadow;
  f3 Ld, shO, shD;
  float shDist;
  bool fromDielectric;
  // An env-map light sample taken ahead of the hit state, so that its
  // texel loads overlap those of the shading record: valid when the light
  // selection draw is reached with seed preSeed (checked where it is used).
  bool preEnv;
  uint32_t preSeed;
  f3 preL, preRad;
  float prePdf;
};

// sampleLights, rchit_layouts.glsl:118-167
__device__ f3 sampleLights(PathCtx& c, f3 pos, f3 nrm, bool& visible, LSample& l) {
  const GpuPushConstantRaytrace& pc = c.sp->pc;
  bool allowDoubleSide = false;
  f3 radiance = mk1(0.f);
  const bool sky = c.sc->sky->inUse == 1;
  bool hasEnv = pc.hasEnvMap == 1 || sky, hasLight = pc.numLights > 0;
  float envSel, anaSel;
  if (hasEnv && hasLight) envSel = anaSel = 0.5f;
  else if (hasEnv) envSel = 1.f, anaSel = 0.f;
  else if (hasLight) envSel = 0.f, anaSel = 1.f;
  else envSel = anaSel = 0.f;
  l.d = l.n = mk1(0.f);
  l.dist = l.pdf = 0.f;
  l.flags = 0;
  const uint32_t selSeed = c.seed;
  float sel = rnd(c.seed);
  if (sel < envSel) {
    l.flags = kLightArea;
    l.dist = kInf;
    l.n = -makeNormal(c.rd);
    float r0 = rnd(c.seed), r1 = rnd(c.seed);
    f3 Ldir, rad;
    float pdf;
    if (sky) {  // sampleSunSky, sample_light.glsl:137-141
      Ldir = uniformSphere(r0, r1);
      pdf = kInv4Pi;
      rad = skyEval(*c.sc->sky, Ldir);
    } else if (c.preEnv && c.preSeed == selSeed) {  // the same draws: the sample taken ahead
      rad = c.preRad;
      Ldir = c.preL;
      pdf = c.prePdf;
    } else {
      rad = sampleEnv(*c.sc, r0, r1, c.sp->cam.envTransform, pc.envMapResolution[0], pc.envMapResolution[1],
                      pc.envMapIntensity, Ldir, pdf);
    }
    l.d = Ldir;
    l.pdf = pdf;
    radiance = rad / envSel;
    allowDoubleSide = true;
  } else if (sel < envSel + anaSel) {
    int li = min(1 + int(rnd(c.seed) * float(pc.numLights)), pc.numLights);
    const GpuLight& light = c.sc->lights[li];
    float r0 = rnd(c.seed), r1 = rnd(c.seed);
    radiance = sampleOneLight(r0, r1, light, pos, l) * float(pc.numLights) / anaSel;
    allowDoubleSide = (light.doubleSide == 1);
  }
  c.shO = offsetRay(pos, nrm);
  c.shD = l.d;
  c.shDist = l.dist;
  visible = (dot(l.d, nrm) > 0.0f && l.pdf > 0.0f);
  visible = visible && (dot(l.n, l.d) < 0.0f || allowDoubleSide);
  return radiance;
}

__device__ __forceinline__ void writeAov(const PathCtx& c, int ch, f3 v, uint32_t pix) {
  if (ch < 0) return;
  size_t plane = size_t(c.sp->width) * c.sp->height;
  c.sp->images[size_t(ch + 1) * plane + pix] = f4(v, 1.f);
}

// hitLight, brdf_lambertian.rchit:44-68
__device__ void hitLight(PathCtx& c, int lightId, f3 hitPos) {
  const GpuLight& light = c.sc->lights[lightId];
  f3 dir = makeNormal(c.rd);
  f3 ln = makeNormal(cross(ld3(light.u), ld3(light.v)));
  float proj = dot(ln, dir);
  c.stop = true;
  if (proj > 0.0f && light.doubleSide == 0) return;
  float mis = 1.0f;
  if (isNonSpecular(c.bflags) && c.depth != 1) {
    float lightDist = length(hitPos - c.ro);
    float distSquare = lightDist * lightDist;
    float lightPdf = distSquare / (light.area * fabsf(proj) + kEps);
    mis = powerHeuristic(c.bpdf, lightPdf);
  }
  c.L = c.L + c.T * ld3(light.radiance) * mis;
}

__device__ __forceinline__ void nextRay(PathCtx& c, f3 pos, f3 offN, f3 d, f3 weightOverPdf, float pdf,
                                        uint32_t flags) {
  c.bpdf = pdf;
  c.bflags = flags;
  c.ro = offsetRay(pos, offN);
  c.rd = d;
  c.T = c.T * weightOverPdf;
}

// brdf_pbr_metalness_roughness.rchit:57-128
__device__ f3 pbrEval(f3 L, f3 V, f3 N, f3 X, f3 Y, f3 albedo, float ax, float ay, float metal, float eta) {
  f3 weight = mk1(0.f);
  float NdotL = dot(N, L), NdotV = dot(N, V);
  if (NdotL <= 0.0f || NdotV <= 0.0f) return weight;
  f3 H = makeNormal(L + V);
  float HdotN = dot(H, N), HdotX = dot(H, X), HdotY = dot(H, Y), HdotV = dot(H, V);
  float F0 = sqr((eta - 1.0f) / (eta + 1.0f));
  f3 f0 = mk1(F0) * (1.0f - metal) + albedo * metal;
  float pw = powf(clampf(1.0f - HdotV, 0.0f, 1.0f), 5.0f);
  f3 Fs = f0 + (mk1(1.f) - f0) * pw;
  float Ds = dAnisoGGX(HdotN, HdotX, HdotY, ax, ay);
  float Gs = g1GGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
  Gs *= g1GGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
  f3 diffuse = albedo * (1.0f - metal) * kInvPi;
  f3 specular = Fs * Ds * Gs;
  weight = weight + diffuse;
  weight = weight + specular;
  weight = weight * NdotL;
  return weight;
}


// brdf_rough_plastic.rchit:82-136 (eval, pdf) with the signatures' parameter order.
__device__ f3 rpEval(f3 L, f3 V, f3 N, f3 X, f3 Y, f3 kd, f3 ks, float eta, float fdrInt, float invEta2, float ax,
                     float ay, uint32_t flags) {
  f3 weight = mk1(0.f);
  float NdotL = dot(L, N), NdotV = dot(V, N);
  if (((flags & kLightArea) == 0) || NdotL < 0.0f || NdotV < 0.0f) return weight;
  {
    f3 H = makeNormal(L + V);
    float Fs = fresnelDielectricExt(dot(H, V), eta);
    float Gs = g1GGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
    Gs *= g1GGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
    float Ds = dAnisoGGX(dot(H, N), dot(H, X), dot(H, Y), ax, ay);
    weight = weight + ks * Fs * Gs * Ds * NdotL;
  }
  {
    float Fo = fresnelDielectricExt(NdotV, eta);
    float Fi = fresnelDielectricExt(NdotL, eta);
    f3 diff = kd / (mk1(1.f) - kd * fdrInt);
    weight = weight + (1.0f - Fi) * (1.0f - Fo) * diff * invEta2 * kInvPi * NdotL;
  }
  return weight;
}
__device__ float rpPdf(f3 L, f3 V, f3 N, f3 X, f3 Y, float ax, float ay, float eta, float ssw, uint32_t flags) {
  float pdf = 0.0f;
  float NdotL = dot(L, N), NdotV = dot(V, N);
  if (NdotL < 0.0f || NdotV < 0.0f || ((flags & kLightArea) == 0)) return pdf;
  float Fo = fresnelDielectricExt(NdotV, eta);
  float sub = ssw * (1.0f - Fo), spec = Fo;
  float ps = spec / (spec + sub);
  f3 wi = toLocal(X, Y, N, L), wo = toLocal(X, Y, N, V);
  f3 wh = makeNormal(wi + wo);
  pdf += ps * pdfGGX(wh, wo, ax, ay);
  pdf += (1.0f - ps) * cosinePdf(wi.z);
  return pdf;
}

// Conductor Fresnel after Shirley (brdf_conductor.rchit:14-34).
__device__ __forceinline__ float conductorR(float eta, float k, float cosThetaI) {
  float cosThetaISq = cosThetaI * cosThetaI;
  float sinThetaISq = fmaxf(1.0f - cosThetaISq, 0.0f);
  float sinThetaIQu = sinThetaISq * sinThetaISq;
  float innerTerm = eta * eta - k * k - sinThetaISq;
  float aSqPlusBSq = sqrtf(fmaxf(innerTerm * innerTerm + 4.0f * eta * eta * k * k, 0.0f));
  float a = sqrtf(fmaxf((aSqPlusBSq + innerTerm) * 0.5f, 0.0f));
  float Rs = ((aSqPlusBSq + cosThetaISq) - (2.0f * a * cosThetaI)) / ((aSqPlusBSq + cosThetaISq) + (2.0f * a * cosThetaI));
  float Rp = ((cosThetaISq * aSqPlusBSq + sinThetaIQu) - (2.0f * a * cosThetaI * sinThetaISq)) /
             ((cosThetaISq * aSqPlusBSq + sinThetaIQu) + (2.0f * a * cosThetaI * sinThetaISq));
  return 0.5f * (Rs + Rs * Rp);
}
__device__ __forceinline__ f3 conductorR3(f3 eta, f3 k, float c) {
  return f3{conductorR(eta.x, k.x, c), conductorR(eta.y, k.y, c), conductorR(eta.z, k.z, c)};
}
// kang18's own luminance weights (brdf_kang18.rchit:60-62)
__device__ __forceinline__ float lum18(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// rough conductor eval, brdf_rough_conductor.rchit:72-91
__device__ f3 rcEval(f3 L, f3 V, f3 N, f3 X, f3 Y, f3 kd, f3 eta, f3 k, float ax, float ay) {
  float NdotL = dot(L, N), NdotV = dot(V, N);
  if (NdotL < 0.0f || NdotV < 0.0f) return mk1(0.f);
  f3 H = makeNormal(L + V);
  f3 Fs = conductorR3(eta, k, NdotL);
  float Gs = g1GGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
  Gs *= g1GGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
  float Ds = dAnisoGGX(dot(H, N), dot(H, X), dot(H, Y), ax, ay);
  return mk1(0.f) + kd * Fs * Gs * Ds * NdotL;
}
// phong, brdf_phong.rchit:12-52
__device__ f3 phongEval(f3 L, f3 V, f3 N, f3 kd, f3 ks, float sh) {
  float NdotL = dot(N, L), NdotV = dot(N, V);
  if (NdotL < 0.0f || NdotV < 0.0f) return mk1(0.f);
  f3 H = makeNormal(L + V);
  f3 diffuseLobe = kd * kInvPi;
  f3 specularLobe = ks * powf(fmaxf(dot(H, N), 0.0f), sh) * (sh + 2.0f) * kInv2Pi;
  float db = luminance(kd), sb = luminance(ks);
  float dw = db / (db + sb);
  float sw = 1.0f - dw;
  return (dw * diffuseLobe + sw * specularLobe) * NdotL;
}
__device__ __forceinline__ float phongLobePdf(f3 L, f3 V, f3 N, float sh) {
  f3 H = makeNormal(L + V);
  float NdotH = fmaxf(dot(H, N), 0.0f), VdotH = fmaxf(dot(H, V), 0.0f);
  return (sh + 1.0f) * kInv2Pi * powf(NdotH, sh) * 0.25f / (VdotH + kEps);
}
// kang18, brdf_kang18.rchit:64-124
__device__ f3 kangEval(f3 L, f3 V, f3 N, f3 X, f3 Y, f3 kd, f3 ks, float ax, float ay, float eta) {
  float NdotL = dot(N, L), NdotV = dot(N, V);
  if (NdotL <= 0.0f || NdotV <= 0.0f) return mk1(0.f);
  f3 H = makeNormal(L + V);
  float HdotN = dot(H, N), HdotX = dot(H, X), HdotY = dot(H, Y), HdotV = dot(H, V);
  float F0 = sqr((eta - 1.0f) / (eta + 1.0f));
  float Fs = F0 + (1.0f - F0) * powf(clampf(1.0f - HdotV, 0.0f, 1.0f), 5.0f);
  float Ds = dAnisoGGX(HdotN, HdotX, HdotY, ax, ay);
  float Gs = g1GGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
  Gs *= g1GGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
  f3 weight = mk1(0.f);
  weight = weight + kd * kInvPi;
  weight = weight + ks * Fs * Ds * Gs;
  return weight * NdotL;
}

// Disney BRDF, brdf_disney.rchit:31-405: Burley diffuse with the
// Hanrahan-Krueger subsurface blend and sheen, anisotropic GGX specular
// sampled through visible normals, GTR1 clearcoat.  GLSL mix(x, y, a) is
// x (1 - a) + y a.  The lobe parameters not held in registers (subsurface,
// tints, sheen, clearcoat) are read from the material record.
struct DsnMat {
  f3 base;
  float metallic, rough, ax, ay;  // rough = max(roughness^2, 0.001)
  const GpuMaterial* m;
};
__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ f3 mix3(f3 x, f3 y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float dsnSchlick(float u) {
  float x = clampf(1.0f - u, 0.0f, 1.0f);
  float x2 = x * x;
  return x2 * x2 * x;
}
__device__ __forceinline__ float dsnDielF(float ci, float eta) {
  float st2 = eta * eta * (1.0f - ci * ci);
  if (st2 > 1.0f) return 1.0f;
  float ct = sqrtf(fmaxf(1.0f - st2, 0.0f));
  float rs = (eta * ct - ci) / (eta * ct + ci);
  float rp = (eta * ci - ct) / (eta * ci + ct);
  return 0.5f * (rs * rs + rp * rp);
}
__device__ __forceinline__ float dsnFresnel(float metallic, float eta, float LdotH, float VdotH) {
  return mixf(dsnDielF(fabsf(VdotH), eta), dsnSchlick(LdotH), metallic);
}
__device__ __forceinline__ float dsnSmithG(float NdotV, float alphaG) {
  float a = alphaG * alphaG, b = NdotV * NdotV;
  return (2.0f * NdotV) / (NdotV + sqrtf(a + b - a * b));
}
__device__ __forceinline__ float dsnSmithGAniso(float NdotV, float VdotX, float VdotY, float ax, float ay) {
  float a = VdotX * ax, b = VdotY * ay, c = NdotV;
  return (2.0f * NdotV) / (NdotV + sqrtf(a * a + b * b + c * c));
}
__device__ __forceinline__ float dsnGTR1(float NdotH, float a) {
  if (a >= 1.0f) return kInvPi;
  float a2 = a * a;
  float t = 1.0f + (a2 - 1.0f) * NdotH * NdotH;
  return (a2 - 1.0f) / (kPi * logf(a2) * t);
}
__device__ __forceinline__ float dsnGTR2Aniso(float NdotH, float HdotX, float HdotY, float ax, float ay) {
  float a = HdotX / ax, b = HdotY / ay;
  float c = a * a + b * b + NdotH * NdotH;
  return 1.0f / (kPi * ax * ay * c * c);
}
__device__ __forceinline__ float dsnCcRough(const DsnMat& d) { return mixf(0.1f, 0.001f, d.m->clearcoatGloss); }
__device__ __forceinline__ f3 dsnEvalDiffuse(const DsnMat& d, f3 csheen, f3 V, f3 L, f3 H, float& pdf) {
  pdf = 0.0f;
  if (L.z <= 0.0f) return mk1(0.f);
  float FL = dsnSchlick(L.z), FV = dsnSchlick(V.z), FH = dsnSchlick(dot(L, H));
  float LH = dot(L, H);
  float Fd90 = 0.5f + 2.0f * LH * LH * d.rough;
  float Fd = mixf(1.0f, Fd90, FL) * mixf(1.0f, Fd90, FV);
  float Fss90 = LH * LH * d.rough;
  float Fss = mixf(1.0f, Fss90, FL) * mixf(1.0f, Fss90, FV);
  float ss = 1.25f * (Fss * (1.0f / (L.z + V.z) - 0.5f) + 0.5f);
  f3 Fsheen = FH * d.m->sheen * csheen;
  pdf = L.z * kInvPi;
  return (1.0f - d.metallic) * (kInvPi * mixf(Fd, ss, d.m->subsurface) * d.base + Fsheen);
}
__device__ __forceinline__ f3 dsnEvalSpec(const DsnMat& d, float eta, f3 specCol, f3 V, f3 L, f3 H, float& pdf) {
  pdf = 0.0f;
  if (L.z <= 0.0f) return mk1(0.f);
  float FM = dsnFresnel(d.metallic, eta, dot(L, H), dot(V, H));
  f3 F = mix3(specCol, mk1(1.f), FM);
  float D = dsnGTR2Aniso(H.z, H.x, H.y, d.ax, d.ay);
  float G1 = dsnSmithGAniso(fabsf(V.z), V.x, V.y, d.ax, d.ay);
  float G2 = G1 * dsnSmithGAniso(fabsf(L.z), L.x, L.y, d.ax, d.ay);
  pdf = G1 * D / (4.0f * V.z);
  return F * D * G2 / (4.0f * L.z * V.z);
}
__device__ __forceinline__ f3 dsnEvalClearcoat(const DsnMat& d, f3 V, f3 L, f3 H, float& pdf) {
  pdf = 0.0f;
  if (L.z <= 0.0f) return mk1(0.f);
  float FH = dsnDielF(dot(V, H), 1.0f / 1.5f);
  float F = mixf(0.04f, 1.0f, FH);
  float D = dsnGTR1(H.z, dsnCcRough(d));
  float G = dsnSmithG(L.z, 0.25f) * dsnSmithG(V.z, 0.25f);
  float jac = 1.0f / (4.0f * dot(V, H));
  pdf = D * H.z * jac;
  return mk1(0.25f) * d.m->clearcoat * F * D * G / (4.0f * L.z * V.z);
}
__device__ __forceinline__ void dsnColors(const DsnMat& d, float eta, f3& specCol, f3& sheenCol) {
  float lum = lum18(d.base);
  f3 ctint = lum > 0.0f ? d.base / lum : mk1(1.f);
  float F0 = (1.0f - eta) / (1.0f + eta);
  specCol = mix3(F0 * F0 * mix3(mk1(1.f), ctint, d.m->specularTint), d.base, d.metallic);
  sheenCol = mix3(mk1(1.f), ctint, d.m->sheenTint);
}
__device__ __forceinline__ void dsnLobes(const DsnMat& d, f3 specCol, float approxF, float& dw, float& sw, float& cw) {
  dw = lum18(d.base) * (1.0f - d.metallic);
  sw = lum18(mix3(specCol, mk1(1.f), approxF));
  cw = 0.25f * d.m->clearcoat * (1.0f - d.metallic);
  float total = dw + sw + cw;
  dw /= total;
  sw /= total;
  cw /= total;
}
// eval (EArea), brdf_disney.rchit:242-296
__device__ f3 dsnEval(f3 Lw, f3 Vw, f3 N, f3 X, f3 Y, const DsnMat& d, float eta, float& bsdfPdf) {
  bsdfPdf = 0.0f;
  f3 V = toLocal(X, Y, N, Vw), L = toLocal(X, Y, N, Lw);
  if (L.z <= 0.0f || V.z <= 0.0f) return mk1(0.f);
  f3 H = makeNormal(L + V);
  if (H.z < 0.0f) H = -H;
  f3 specCol, sheenCol;
  dsnColors(d, eta, specCol, sheenCol);
  float dw, sw, cw;
  dsnLobes(d, specCol, dsnFresnel(d.metallic, eta, dot(L, H), dot(V, H)), dw, sw, cw);
  f3 w = mk1(0.f);
  float pdf;
  if (dw > 0.0f) {
    w = w + dsnEvalDiffuse(d, sheenCol, V, L, H, pdf);
    bsdfPdf += pdf * dw;
  }
  if (sw > 0.0f) {
    w = w + dsnEvalSpec(d, eta, specCol, V, L, H, pdf);
    bsdfPdf += pdf * sw;
  }
  if (cw > 0.0f) {
    w = w + dsnEvalClearcoat(d, V, L, H, pdf);
    bsdfPdf += pdf * cw;
  }
  return w * L.z;
}
// sampleBsdf, brdf_disney.rchit:298-363.  SampleGTR1 draws phi and
// cos theta from the same r1 (r2 unused), as written.  The weight is
// f |dot(N, L)| with N in world space and L local, as written.
__device__ f3 dsnSample(float r1, float r2, f3 Vw, f3 N, f3 X, f3 Y, const DsnMat& d, float eta, f3& dir,
                        float& pdf, uint32_t& flags) {
  pdf = 0.0f;
  f3 V = toLocal(X, Y, N, Vw), L, f;
  f3 specCol, sheenCol;
  dsnColors(d, eta, specCol, sheenCol);
  float dw, sw, cw;
  dsnLobes(d, specCol, dsnFresnel(d.metallic, eta, V.z, V.z), dw, sw, cw);
  float cdf0 = dw, cdf1 = cdf0 + cw;
  if (r1 < cdf0) {
    r1 /= cdf0;
    L = cosineHemisphere(r1, r2);
    f = dsnEvalDiffuse(d, sheenCol, V, L, makeNormal(L + V), pdf);
    pdf *= dw;
    flags = kDiffuseRefl;
  } else if (r1 < cdf1) {
    r1 = (r1 - cdf0) / (cdf1 - cdf0);
    float a = fmaxf(0.001f, dsnCcRough(d));
    float a2 = a * a;
    float phi = r1 * kTwoPi;
    float ct = sqrtf((1.0f - powf(a2, 1.0f - r1)) / (1.0f - a2));
    float st = clampf(sqrtf(1.0f - ct * ct), 0.0f, 1.0f);
    f3 H = f3{st * cosf(phi), st * sinf(phi), ct};
    if (H.z < 0.0f) H = -H;
    L = makeNormal(reflect(-V, H));
    f = dsnEvalClearcoat(d, V, L, H, pdf);
    pdf *= cw;
    flags = kGlossyRefl;
  } else {
    r1 = (r1 - cdf1) / (1.0f - cdf1);
    // visible-normal sampling (Heitz, JCGT 2018), brdf_disney.rchit:96-115
    f3 Vh = makeNormal(f3{d.ax * V.x, d.ay * V.y, V.z});
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    f3 T1 = lensq > 0.0f ? f3{-Vh.y, Vh.x, 0.0f} * (1.0f / sqrtf(lensq)) : f3{1.0f, 0.0f, 0.0f};
    f3 T2 = cross(Vh, T1);
    float r = sqrtf(r1), phi = 2.0f * kPi * r2;
    float t1 = r * cosf(phi), t2 = r * sinf(phi);
    float sv = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - sv) * sqrtf(1.0f - t1 * t1) + sv * t2;
    f3 Nh = t1 * T1 + t2 * T2 + sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2)) * Vh;
    f3 H = makeNormal(f3{d.ax * Nh.x, d.ay * Nh.y, fmaxf(0.0f, Nh.z)});
    if (H.z < 0.0f) H = -H;
    L = makeNormal(reflect(-V, H));
    f = dsnEvalSpec(d, eta, specCol, V, L, H, pdf);
    pdf *= sw;
    flags = kGlossyRefl;
  }
  dir = toWorld(X, Y, N, L);
  return f * fabsf(dot(N, L));
}

// Surface plugins.  Every closest-hit shader of the set runs the same three
// steps -- material inputs (textures, normal map, AOVs, opacity draws), one
// light sample through sampleLights, then its BSDF value toward the light
// and its BSDF sample -- and the light sample sits at the same point of the
// RNG stream in each.  The light sample is therefore taken once for the
// whole wave instead of once per plugin present in it.
//   brdf_lambertian.rchit:70-151, brdf_pbr_metalness_roughness.rchit:131-219,
//   brdf_rough_plastic.rchit:191-269, bsdf_dielectric.rchit:74-138,
//   brdf_mirror.rchit:39-101, brdf_conductor.rchit:89-155,
//   brdf_rough_conductor.rchit:147-220, brdf_plastic.rchit:133-204,
//   brdf_phong.rchit:112-193, brdf_kang18.rchit:141-249,
//   brdf_disney.rchit:365-477
//
// The plugins are compiled as two groups (GROUP2 = false: lambertian, pbr,
// rough plastic, dielectric; true: mirror, conductor, rough conductor,
// plastic, phong, kang18, disney) so that each shading kernel keeps the
// register budget of one group.

template <bool GROUP2>
__device__ void shadeSurface(PathCtx& c, Hit& h, const GpuMaterial& m, uint32_t pix, bool aov) {
  const GpuPushConstantRaytrace& pc = c.sp->pc;
  const SceneView& sc = *c.sc;
  const uint32_t type = m.type;
  const bool lambert = !GROUP2 && type == MaterialTypeBrdfLambertian,
             pbr = !GROUP2 && type == MaterialTypeBrdfPbrMetalnessRoughness,
             plastic = !GROUP2 && type == MaterialTypeBrdfRoughPlastic,
             dielectric = !GROUP2 && type == MaterialTypeBsdfDielectric,
             mirror = GROUP2 && type == MaterialTypeBrdfMirror, conductor = GROUP2 && type == MaterialTypeBrdfConductor,
             rcond = GROUP2 && type == MaterialTypeBrdfRoughConductor, smooth = GROUP2 && type == MaterialTypeBrdfPlastic,
             phong = GROUP2 && type == MaterialTypeBrdfPhong, kang = GROUP2 && type == MaterialTypeBrdfKang18,
             disney = GROUP2 && type == MaterialTypeBrdfDisney;
  // NEE is identically zero for the delta lobes (their eval tests EDelta but
  // is called with EArea); their shadow ray is counted, not traced.
  const bool zeroNee = dielectric || mirror || conductor;
  // (1) material inputs
  f3 kd = ld3(m.diffuse), ks = mk1(1.f);
  float metal = 0.f, ax = 0.f, ay = 0.f, eta = m.ior, fdrInt = 0.f, invEta2 = 0.f, ssw = 0.f;
  f3 ceta = mk1(0.f), ck = mk1(0.f);
  if (!dielectric && m.diffuseTextureId >= 0) kd = xyz(tex2D(sc.textures[m.diffuseTextureId], h.uvx, h.uvy));
  if (pbr || disney) {
    metal = m.metalness;
    float rough = m.roughness;
    if (m.metalnessTextureId >= 0) metal = tex2D(sc.textures[m.metalnessTextureId], h.uvx, h.uvy).x;
    if (m.roughnessTextureId >= 0) rough = tex2D(sc.textures[m.roughnessTextureId], h.uvx, h.uvy).x;
    ax = ay = fmaxf(sqr(rough), 0.001f);
    if (disney) {  // brdf_disney.rchit:419-421
      float aspect = sqrtf(1.0f - m.anisotropic * 0.9f);
      fdrInt = ax;  // max(roughness^2, 0.001)
      ax = fmaxf(0.001f, rough * rough / aspect);
      ay = fmaxf(0.001f, rough * rough * aspect);
    }
  } else if (plastic || rcond || kang) {
    float alx = m.anisoAlpha[0], aly = m.anisoAlpha[1];
    if (m.roughnessTextureId >= 0) {
      float4 t = tex2D(sc.textures[m.roughnessTextureId], h.uvx, h.uvy);
      alx = t.x;
      aly = t.y;
    }
    ax = fmaxf(plastic ? kEps : alx, plastic ? alx : kEps);
    ay = fmaxf(plastic ? kEps : aly, plastic ? aly : kEps);
    if (plastic) {
      fdrInt = m.radiance[0];
      invEta2 = 1.0f / (eta * eta);
      ssw = luminance(kd);
    }
  } else if (smooth) {
    fdrInt = m.radiance[0];
    float dAvg = luminance(kd), sAvg = luminance(mk1(1.f));
    ssw = sAvg / (dAvg + sAvg);
    invEta2 = 1.0f / (eta * eta);
  }
  if (conductor || rcond) {
    ceta = ld3(m.radiance);
    ck = ld3(m.radianceFactor);
  }
  if (phong) {
    ks = ld3(m.rhoSpec);
    metal = m.specular;  // shininess
  }
  if (kang) {
    ks = ld3(m.rhoSpec);
    if (m.metalnessTextureId >= 0) ks = xyz(tex2D(sc.textures[m.metalnessTextureId], h.uvx, h.uvy));
    float opacity = m.metalness;
    if (m.opacityTextureId >= 0) opacity = tex2D(sc.textures[m.opacityTextureId], h.uvx, h.uvy).x;
    const float* q = sc.inst[h.inst].w2o;
    auto toWorldN = [&](f3 n) {  // (n * gl_WorldToObjectEXT).xyz
      return f3{n.x * q[0] + n.y * q[4] + n.z * q[8], n.x * q[1] + n.y * q[5] + n.z * q[9],
                n.x * q[2] + n.y * q[6] + n.z * q[10]};
    };
    if (m.normalTextureId >= 0) {
      f3 cn = xyz(tex2D(sc.textures[m.normalTextureId], h.uvx, h.uvy));
      h.N = makeNormal(toWorldN(2.0f * cn - mk1(1.f)));
      if (pc.useFaceNormal == 1) h.N = mk1(0.f);
      basis(h.N, h.X, h.Y);
      h.ffN = dot(h.N, h.V) > 0.0f ? h.N : -h.N;
    }
    if (m.tangentTextureId >= 0) {
      f3 ct = xyz(tex2D(sc.textures[m.tangentTextureId], h.uvx, h.uvy));
      h.X = makeNormal(toWorldN(2.0f * ct - mk1(1.f)));
    }
    h.Y = makeNormal(cross(h.N, h.X));
    h.X = makeNormal(cross(h.Y, h.N));
    h.ffN = dot(h.N, h.V) > 0.0f ? h.N : -h.N;
    if (rnd(c.seed) < opacity) {
      c.ro = offsetRay(h.pos, -h.ffN);
      c.depth--;
      return;
    }
  } else if (m.normalTextureId >= 0) {
    normalMap(sc, h, m.normalTextureId, pc.useFaceNormal);
  }
  if (pbr || disney) {
    float opacity = pbr ? m.specular : m.rhoSpec[0];
    if (m.opacityTextureId >= 0) opacity = tex2D(sc.textures[m.opacityTextureId], h.uvx, h.uvy).x;
    if (rnd(c.seed) < opacity) {
      // pass through: same direction, same depth (pbr:156-159)
      c.ro = offsetRay(h.pos, -h.ffN);
      c.depth--;
      return;
    }
  }
  if (dielectric || disney) eta = dot(h.V, h.N) > 0.0f ? (1.0f / m.ior) : m.ior;
  const DsnMat dm{kd, metal, fdrInt, ax, ay, &m};
  if (aov) {
    if (lambert) {
      writeAov(c, pc.diffuseOutChannel, kd, pix);
      writeAov(c, pc.normalOutChannel, h.N, pix);
      writeAov(c, pc.specularOutChannel, mk1(0.f), pix);
      writeAov(c, pc.tangentOutChannel, h.X, pix);
      writeAov(c, pc.roughnessOutChannel, f3{1.f, 1.f, 0.f}, pix);
      writeAov(c, pc.positionOutChannel, h.pos, pix);
      writeAov(c, pc.uvOutChannel, f3{h.uvx, h.uvy, 1.f}, pix);
    } else if (pbr) {
      writeAov(c, pc.diffuseOutChannel, kd, pix);
      writeAov(c, pc.normalOutChannel, h.ffN, pix);
    } else if (phong || kang) {
      writeAov(c, pc.diffuseOutChannel, kd, pix);
      writeAov(c, pc.normalOutChannel, h.ffN, pix);
      writeAov(c, pc.specularOutChannel, ks, pix);
      writeAov(c, pc.tangentOutChannel, h.X, pix);
      writeAov(c, pc.roughnessOutChannel, kang ? f3{ax, ay, 0.f} : f3{1.f, 1.f, 0.f}, pix);
      writeAov(c, pc.positionOutChannel, h.pos, pix);
      writeAov(c, pc.uvOutChannel, f3{h.uvx, h.uvy, 1.f}, pix);
    }
  }
  // (2) one light sample (sampleLights, rchit_layouts.glsl:118-167)
  bool visible;
  LSample l;
  f3 radiance = sampleLights(c, h.pos, h.ffN, visible, l);
  // (3) BSDF value and pdf toward the light
  f3 Ld = mk1(0.f);
  if (visible && !zeroNee) {
    f3 bw = mk1(0.f);
    float bsdfPdf = 0.0f;
    const f3 N = h.ffN, V = h.V;
    if (lambert) {
      float NdotL = dot(N, l.d), NdotV = dot(N, V);
      if (!(NdotL < 0.0f || NdotV < 0.0f)) bw = kd * kInvPi * NdotL;
      bsdfPdf = (l.flags & kLightArea) ? cosinePdf(dot(N, l.d)) : 0.0f;
    } else if (pbr) {
      bw = pbrEval(l.d, V, N, h.X, h.Y, kd, ax, ay, metal, eta);
      if (l.flags & kLightArea) {
        float NdotL = dot(N, l.d), NdotV = dot(N, V);
        if (!(NdotL <= 0.0f || NdotV <= 0.0f)) {
          f3 H = makeNormal(l.d + V);
          f3 wh = makeNormal(toLocal(h.X, h.Y, N, H));
          f3 wo = makeNormal(toLocal(h.X, h.Y, N, V));
          f3 wi = makeNormal(toLocal(h.X, h.Y, N, l.d));
          bsdfPdf += 0.2f * cosinePdf(wi.z);
          bsdfPdf += (1.0f - 0.2f) * pdfGGX(wh, wo, ax, ay);
        }
      }
    } else if (plastic) {
      // The shader passes (..., ax, ay, invEta2, EArea) into eval's
      // (..., invEta2, ax, ay, flags) and (eta, ax, ay) into pdf's
      // (ax, ay, eta) (brdf_rough_plastic.rchit:234-240).
      bw = rpEval(l.d, V, N, h.X, h.Y, kd, ks, eta, fdrInt, ax, ay, invEta2, kLightArea);
      bsdfPdf = rpPdf(l.d, V, N, h.X, h.Y, eta, ax, ay, ssw, l.flags);
    } else if (rcond) {
      bw = rcEval(l.d, V, N, h.X, h.Y, kd, ceta, ck, ax, ay);
      float NdotL = dot(l.d, N), NdotV = dot(V, N);
      if (!(NdotL < 0.0f || NdotV < 0.0f) && (l.flags & kLightArea)) {
        f3 wi = toLocal(h.X, h.Y, N, l.d), wo = toLocal(h.X, h.Y, N, V);
        bsdfPdf = pdfGGX(makeNormal(wi + wo), wo, ax, ay);
      }
    } else if (smooth) {
      float NdotL = dot(l.d, N), NdotV = dot(V, N);
      if (!(NdotL < 0.0f || NdotV < 0.0f)) {
        float Fo = fresnelDielectricExt(NdotV, eta), Fi = fresnelDielectricExt(NdotL, eta);
        f3 diff = kd / (mk1(1.f) - kd * fdrInt);
        bw = (1.0f - Fi) * (1.0f - Fo) * diff * invEta2 * kInvPi * NdotL;
        float ps = (Fo * ssw) / (Fo * ssw + (1.0f - Fo) * (1.0f - ssw));
        if (l.flags & kLightDelta) bsdfPdf = fabsf(dot(reflect(-V, N), l.d) - 1.0f) < kEps ? ps : 0.0f;
        else if (l.flags & kLightArea) bsdfPdf = NdotL * (1.0f - ps);  // no 1/pi, as written
      }
    } else if (phong) {
      bw = phongEval(l.d, V, N, kd, ks, metal);
      if (l.flags & kLightArea) {
        float db = luminance(kd), sb = luminance(ks);
        float dw = db / (db + sb);
        float sw = 1.0f - dw;
        bsdfPdf = dw * cosinePdf(dot(N, l.d)) + sw * phongLobePdf(l.d, V, N, metal);
      }
    } else if (kang) {  // kang18: eval ignores the flags
      bw = kangEval(l.d, V, N, h.X, h.Y, kd, ks, ax, ay, eta);
      float NdotL = dot(N, l.d), NdotV = dot(N, V);
      if ((l.flags & kLightArea) && !(NdotL <= 0.0f || NdotV <= 0.0f)) {
        f3 H = makeNormal(l.d + V);
        f3 wh = makeNormal(toLocal(h.X, h.Y, N, H)), wo = makeNormal(toLocal(h.X, h.Y, N, V)),
           wi = makeNormal(toLocal(h.X, h.Y, N, l.d));
        float pd = lum18(kd), ps = lum18(ks);
        float P = pd / (pd + ps + kEps);
        bsdfPdf += P * cosinePdf(wi.z);
        bsdfPdf += (1.0f - P) * pdfGGX(wh, wo, ax, ay);
      }
    } else if (disney) {  // eval is called with EArea whatever the light
      bw = dsnEval(l.d, V, N, h.X, h.Y, dm, eta, bsdfPdf);
    }
    float mis = powerHeuristic(l.pdf, bsdfPdf);
    Ld = mis * bw * radiance * c.T / (l.pdf + kEps);
  }
  c.Ld = Ld;
  c.shadow = visible && !zeroNee;
  c.fromDielectric = visible && zeroNee;
  // (4) BSDF sample
  if (dielectric) {
    float u = rnd(c.seed);
    f3 V = h.V, N = h.ffN;
    float F = dielectricFresnel(fabsf(dot(V, N)), eta);
    f3 d, w;
    float pdf;
    uint32_t flags;
    if (u < F) {
      d = makeNormal(reflect(-V, N));
      pdf = F;
      flags = kSpecRefl;
      w = F * mk1(1.f);
    } else {
      d = makeNormal(refract(-V, N, eta));
      pdf = 1.0f - F;
      flags = kSpecTrans;
      w = (1.0f - F) * mk1(1.f) * eta * eta;
    }
    if (pdf <= 0.0f || length(w) == 0.0f) {
      c.stop = true;
      return;
    }
    float sg = dot(d, h.N);
    float sgn = sg > 0.0f ? 1.0f : (sg < 0.0f ? -1.0f : 0.0f);
    c.bpdf = pdf;
    c.bflags = flags;
    c.ro = offsetRay(h.pos, sgn * h.N);
    c.rd = d;
    c.T = c.T * (w / (pdf + kEps));
    return;
  }
  if (mirror) {  // sampleBsdf takes no random numbers
    if (length(kd) == 0.0f) {
      c.stop = true;
      return;
    }
    nextRay(c, h.pos, h.ffN, reflect(-h.V, h.ffN), kd / 1.0f, 1.0f, kSpecRefl);
    return;
  }
  float u0 = rnd(c.seed), u1 = rnd(c.seed);
  if (lambert) {
    f3 wi = cosineHemisphere(u0, u1);
    float pdf = cosinePdf(wi.z);
    f3 d = toWorld(h.X, h.Y, h.ffN, wi);
    f3 w = kd * kInvPi * fabsf(wi.z);
    if (pdf <= 0.0f || length(w) == 0.0f) {
      c.stop = true;
      return;
    }
    nextRay(c, h.pos, h.ffN, d, w / pdf, pdf, kDiffuseRefl);
    return;
  }
  f3 d = mk1(0.f), w = mk1(0.f);
  float pdf = 0.0f;
  uint32_t flags = kBsdfNull;
  bool epsInDivisor = true;
  const f3 N = h.ffN, V = h.V, X = h.X, Y = h.Y;
  if (pbr) {
    f3 wo = makeNormal(toLocal(X, Y, N, V));
    f3 wi;
    if (rnd(c.seed) < 0.2f) {
      wi = cosineHemisphere(u0, u1);
      flags = kDiffuseRefl;
      pdf = cosinePdf(wi.z);
    } else {
      wi = sampleGGX(u0, u1, wo, ax, ay);
      f3 wh = makeNormal(wi + wo);
      pdf = pdfGGX(wh, wo, ax, ay);
      flags = kGlossyRefl;
    }
    d = toWorld(X, Y, N, wi);
    w = pbrEval(d, V, N, X, Y, kd, ax, ay, metal, eta) * fabsf(wi.z);
  } else if (plastic) {
    // sampleBsdf, brdf_rough_plastic.rchit:138-189
    float NdotV = dot(V, N);
    if (NdotV > 0.0f) {
      float Fo = fresnelDielectricExt(NdotV, eta);
      float sub = ssw * (1.0f - Fo), spec = Fo;
      float ps = spec / (spec + sub);
      if (u0 < ps) {
        u0 = u0 / ps;
        f3 wo = toLocal(X, Y, N, V);
        f3 wi = sampleGGX(u0, u1, wo, ax, ay);
        f3 Lw = toWorld(X, Y, N, wi);
        d = Lw;
        f3 H = makeNormal(V + Lw);
        f3 wh = makeNormal(wi + wo);
        float NdotL = dot(N, Lw);
        float Fs = fresnelDielectricExt(dot(H, V), eta);
        float Gs = g1GGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
        Gs *= g1GGX(NdotL, dot(Lw, X), dot(Lw, Y), ax, ay);
        float Ds = dAnisoGGX(dot(H, N), dot(H, X), dot(H, Y), ax, ay);
        flags = kGlossyRefl;
        pdf = pdfGGX(wh, wo, ax, ay) * ps;
        w = ks * Fs * Gs * Ds * NdotL;
      } else {
        u0 = (u0 - ps) / (1.0f + kEps - ps);
        f3 wi = cosineHemisphere(u0, u1);
        float Fi = fresnelDielectricExt(wi.z, eta);
        f3 diff = kd / (mk1(1.f) - kd * fdrInt);
        pdf = (1.0f - ps) * cosinePdf(wi.z);
        d = toWorld(X, Y, N, wi);
        flags = kDiffuseRefl;
        w = (1.0f - Fi) * (1.0f - Fo) * invEta2 * diff * kInvPi * fabsf(wi.z);
      }
    }
  } else if (conductor) {  // the two draws are unused
    if (dot(V, N) > 0.0f) {
      d = reflect(-V, N);
      pdf = 1.0f;
      flags = kSpecRefl;
      w = kd * conductorR3(ceta, ck, dot(N, d));
    }
  } else if (rcond) {
    float NdotV = dot(V, N);
    if (NdotV > 0.0f) {
      f3 wo = toLocal(X, Y, N, V);
      f3 wi = sampleGGX(u0, u1, wo, ax, ay);
      f3 L = toWorld(X, Y, N, wi);
      d = L;
      f3 H = makeNormal(V + L);
      f3 wh = makeNormal(wi + wo);
      float NdotL = dot(N, L);
      f3 Fs = conductorR3(ceta, ck, NdotL);
      float Gs = g1GGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
      Gs *= g1GGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
      float Ds = dAnisoGGX(dot(H, N), dot(H, X), dot(H, Y), ax, ay);
      flags = kGlossyRefl;
      pdf = pdfGGX(wh, wo, ax, ay);
      w = kd * Fs * Gs * Ds * NdotL;
    }
  } else if (smooth) {
    float NdotV = dot(V, N);
    if (NdotV > 0.0f) {
      float Fo = fresnelDielectricExt(NdotV, eta);
      float ps = (Fo * ssw) / (Fo * ssw + (1.0f - Fo) * (1.0f - ssw));
      if (u0 < ps) {
        d = makeNormal(reflect(-V, N));
        flags = kSpecRefl;
        pdf = ps;
        w = mk1(1.f) * Fo;
      } else {
        u0 = (u0 - ps) / (1.0f + kEps - ps);
        f3 wi = cosineHemisphere(u0, u1);
        float Fi = fresnelDielectricExt(wi.z, eta);
        f3 diff = kd / (mk1(1.f) - kd * fdrInt);
        pdf = (1.0f - ps) * cosinePdf(wi.z);
        d = toWorld(X, Y, N, wi);
        flags = kDiffuseRefl;
        w = (1.0f - Fi) * (1.0f - Fo) * invEta2 * diff * kInvPi * fabsf(wi.z);
      }
    }
  } else if (phong) {
    epsInDivisor = false;
    float db = luminance(kd), sb = luminance(ks);
    float dw = db / (db + sb);
    float sw = 1.0f - dw;
    if (u0 < dw) {
      u0 /= dw;
      f3 wi = cosineHemisphere(u0, u1);
      pdf = cosinePdf(wi.z);
      d = toWorld(X, Y, N, wi);
      flags = kDiffuseRefl;
    } else {
      u0 = (u0 - dw) / sw;
      float cosTheta = powf(u0, 1.0f / (metal + 1.0f));
      float phi = kTwoPi * u1;
      float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
      f3 wh = f3{sinTheta * sinf(phi), sinTheta * cosf(phi), cosTheta};
      f3 H = toWorld(X, Y, N, wh);
      d = reflect(-V, H);
      pdf = phongLobePdf(d, V, N, metal);
      flags = kGlossyRefl;
    }
    w = phongEval(d, V, N, kd, ks, metal);
  } else if (kang) {
    f3 wo = makeNormal(toLocal(X, Y, N, V));
    float pd = lum18(kd), ps = lum18(ks);
    float P = pd / (pd + ps + kEps);
    f3 wi;
    if (rnd(c.seed) < P) {
      wi = cosineHemisphere(u0, u1);
      flags = kDiffuseRefl;
      pdf = cosinePdf(wi.z);
    } else {
      wi = sampleGGX(u0, u1, wo, ax, ay);
      pdf = pdfGGX(makeNormal(wi + wo), wo, ax, ay);
      flags = kGlossyRefl;
    }
    d = toWorld(X, Y, N, wi);
    w = kangEval(d, V, N, X, Y, kd, ks, ax, ay, eta);
  } else if (disney) {
    epsInDivisor = false;
    w = dsnSample(u0, u1, V, N, X, Y, dm, eta, d, pdf, flags);
  }
  if (pdf <= 0.0f || length(w) == 0.0f) {
    c.stop = true;
    return;
  }
  nextRay(c, h.pos, h.ffN, d, epsInDivisor ? w / (pdf + kEps) : w / pdf, pdf, flags);
}

// One bounce of shading for one queued path: miss, light hit, emissive or
// surface plugin; appends the NEE shadow ray and the continuing path.
// ONE: the batch holds a single shot whose constants come from the kernel
// arguments (scalar loads) instead of the per-shot table.
// MATS says where the plugins read the hit's material from:
//   kMatGlobal  global memory, field by field;
//   kMatStaged  sMat is this thread's 144-byte LDS slot; the material is
//               fetched into it with nine vector loads issued together;
//   kMatTable   sMat holds the scene's whole material table, copied into
//               LDS by the block (scenes with at most kMatTableMax materials).
enum MatSource { kMatGlobal = 0, kMatStaged = 1, kMatTable = 2 };
constexpr uint32_t kMatTableMax = 128;
constexpr int kShadeKeys = 16;  // 0 miss, 1 emitter, 2 + material type; kShadeKeys - 1 = not shaded here
constexpr int kShadeBlock = 512;
constexpr int kShadeWaves = kShadeBlock / 64;

// Block-wide appends to the shadow-ray stream and the next queue, plus the
// delta-lobe NEE count: one atomic per counter and block instead of one per
// wave (device-scope atomics on one address are resolved beyond the XCD's
// L2 and serialise). Every thread of the block must call it.
struct BlockAppend {
  uint32_t n[3][kShadeWaves];
  uint32_t base[2];
};
__device__ __forceinline__ void blockAppend(BlockAppend& s, uint32_t* ctrA, bool pa, uint32_t* ctrB, bool pb,
                                            unsigned long long* cnt, bool pn, int& ia, int& ib) {
  const uint64_t ma = __ballot(pa), mb = __ballot(pb), mn = __ballot(pn);
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  if (lane == 0) {
    s.n[0][wave] = uint32_t(__popcll(ma));
    s.n[1][wave] = uint32_t(__popcll(mb));
    s.n[2][wave] = uint32_t(__popcll(mn));
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const uint32_t k = threadIdx.x;
    uint32_t run = 0;
    for (int v = 0; v < kShadeWaves; ++v) {
      const uint32_t c = s.n[k][v];
      s.n[k][v] = run;
      run += c;
    }
    if (k == 2) {
      if (run) atomicAdd(cnt, (unsigned long long)run);
    } else {
      s.base[k] = run ? atomicAdd(k == 0 ? ctrA : ctrB, run) : 0u;
    }
  }
  __syncthreads();
  const uint64_t below = lane ? ((~0ull) >> (64 - lane)) : 0ull;
  ia = pa ? int(s.base[0] + s.n[0][wave] + uint32_t(__popcll(ma & below))) : -1;
  ib = pb ? int(s.base[1] + s.n[1][wave] + uint32_t(__popcll(mb & below))) : -1;
}

template <bool ONE, bool GROUP2, int MATS, bool BLOCK>
__device__ __forceinline__ void shadeOne(const SceneView& sc, const BatchDesc& b, const Wavefront& w, int cur,
                                         int slot, float4* sMat, BlockAppend* sApp) {
  const bool active = slot >= 0;
  int pid = -1;
  PathCtx c;
  c.stop = true;
  c.shadow = false;
  c.fromDielectric = false;
  uint4 info = make_uint4(0u, 0u, 0u, 0u);
  float4 rad = make_float4(0.f, 0.f, 0.f, 0.f);
  bool cont = false;
  if (active) {
    const float4 d4 = w.rayD[cur][slot];
    pid = __float_as_int(d4.w);
    info = w.info[cur][slot];
    uint32_t shot = info.w & 0xffffu;
    const ShotParams& sp = ONE ? b.one : b.shotTab[shot];
    const GpuPushConstantRaytrace& pc = sp.pc;
    // a depth-1 path (a camera path, or one that passed through a cut-out
    // surface, which keeps its throughput) has throughput (1, 1, 1) and pdf
    // 0; raygen stores no record for it
    const float4 thr = (info.y & 0xffffu) == 1u ? make_float4(1.f, 1.f, 1.f, 0.f) : w.thr[cur][slot];
    rad = w.rad[pid];
    c.sc = &sc;
    c.sp = &sp;
    c.seed = info.x;
    c.depth = info.y & 0xffffu;
    c.bflags = info.y >> 16;
    c.rd = xyz(d4);  // the origin is read only for light hits (MIS distance)
    c.T = xyz(thr);
    c.bpdf = thr.w;
    c.L = xyz(rad);
    c.stop = false;
    c.shadow = false;
    c.fromDielectric = false;
    c.preEnv = false;
    c.Ld = mk1(0.f);
    int4 hi = w.hitI[slot];
    uint32_t pix = info.z;
    bool frame0 = (info.w >> 16) & 1u;
    if (hi.x < 0) {
      // raytrace.default.rmiss:24-55
      c.stop = true;
      const bool sky = sc.sky->inUse == 1;
      const bool envMap = !sky && sc.hasEnv && pc.hasEnvMap == 1;
      float envPdf = kInv4Pi;
      f3 env = sky      ? skyEval(*sc.sky, c.rd)
               : envMap ? evalEnvPdf(sc, sp.cam.envTransform, pc.envMapIntensity, pc.envMapResolution[0],
                                     pc.envMapResolution[1], c.rd, envPdf)
                        : ld3(pc.bgColor);
      float mis = 1.0f;
      if (c.depth != 1 && isNonSpecular(c.bflags)) mis = powerHeuristic(c.bpdf, envPdf);
      c.L = c.L + c.T * env * mis;
    } else {
      // Surfaces draw their light sample after a fixed number of draws: one
      // (the opacity draw) for pbr, disney and kang18, none for the others.
      // With an env map, the env branch of that sample depends on nothing
      // but those draws, so it is taken here; sampleLights uses it only if
      // its selection draw starts from the same seed.
      const uint32_t sw = uint32_t(hi.y);
      if (!(sw & kShadeLight) && sc.hasEnv && pc.hasEnvMap == 1 && sc.sky->inUse != 1) {
        const uint32_t t = (sw >> kShadeTypeShift) & kShadeTypeMask;
        uint32_t s = c.seed;
        if (t == MaterialTypeBrdfPbrMetalnessRoughness || t == MaterialTypeBrdfDisney || t == MaterialTypeBrdfKang18)
          (void)rnd(s);
        c.preSeed = s;
        if (rnd(s) < (pc.numLights > 0 ? 0.5f : 1.0f)) {
          const float r0 = rnd(s), r1 = rnd(s);
          c.preRad = sampleEnv(sc, r0, r1, sp.cam.envTransform, pc.envMapResolution[0], pc.envMapResolution[1],
                               pc.envMapIntensity, c.preL, c.prePdf);
          c.preEnv = true;
        }
      }
      float4 hf = w.hitF[slot];
      Hit h;
      hitState(sc, hi.x, uint32_t(hi.y), hi.w, hf.y, hf.z, c.rd, pc.useFaceNormal, h);
      bool aov = frame0 && c.depth == 1;
      if (aov && sp.depthChannel >= 0) writeAov(c, sp.depthChannel, mk1(hf.x), pix);
      if (h.lightId >= 0) {
        c.ro = xyz(w.rayO[cur][slot]);
        hitLight(c, h.lightId, h.pos);
      } else {
        if (MATS == kMatStaged) {
          const float4* src = reinterpret_cast<const float4*>(sc.mats + h.mat);
          float4 mv[9];
  #pragma unroll
          for (int k = 0; k < 9; ++k) mv[k] = src[k];
  #pragma unroll
          for (int k = 0; k < 9; ++k) sMat[k] = mv[k];
        }
        const GpuMaterial& m = MATS == kMatGlobal   ? sc.mats[h.mat].m
                               : MATS == kMatTable ? *reinterpret_cast<const GpuMaterial*>(sMat + 9 * h.mat)
                                                   : *reinterpret_cast<const GpuMaterial*>(sMat);
        uint32_t type = m.type;
        if (type == MaterialTypeBrdfEmissive) {
          // brdf_emissive.rchit:12-26
          c.stop = true;
          f3 e = ld3(m.radiance);
          if (m.radianceTextureId >= 0) e = ld3(m.radianceFactor) * xyz(tex2D(sc.textures[m.radianceTextureId], h.uvx, h.uvy));
          if (pc.ignoreEmissive == 0) c.L = c.L + e * c.T;
        } else if (type < MaterialTypeNum) {
          shadeSurface<GROUP2>(c, h, m, pix, aov);
        } else {
          c.stop = true;
        }
      }
    }
    uint32_t nd = c.depth + 1;
    cont = !c.stop && int32_t(nd) <= pc.maxPathDepth;
    c.depth = c.stop ? c.depth : nd;
  }
  // Appends: the NEE shadow ray (maxDist = dist - 2*EPS, rgen:117-122), the
  // continuing path, and the count of delta-lobe NEE rays (counted, not
  // traced; the extension and shadow ray counts are added per bounce by
  // kStatRays).
  int si, ni;
  if (BLOCK) {
    blockAppend(*sApp, &w.ctr[2], c.shadow, &w.ctr[1 - cur], cont, &w.cnt64[6], c.fromDielectric, si, ni);
  } else {
    waveCount64(&w.cnt64[6], c.fromDielectric);
    si = waveAppend(&w.ctr[2], c.shadow);
    ni = waveAppend(&w.ctr[1 - cur], cont);
  }
  if (si >= 0) {
    w.shO[si] = f4(c.shO, c.shDist - 2.0f * kEps);
    w.shD[si] = f4(c.shD, __int_as_float(pid));
    w.ld[si] = f4(c.Ld, 0.f);
  }
  if (ni >= 0) {  // the continuing path, in its slot of the next queue
    w.rayO[1 - cur][ni] = f4(c.ro, 0.f);
    w.rayD[1 - cur][ni] = f4(c.rd, __int_as_float(pid));
    w.thr[1 - cur][ni] = f4(c.T, c.bpdf);
    w.info[1 - cur][ni] = make_uint4(c.seed, c.depth | (c.bflags << 16), info.z, info.w);
  }
  if (active) w.rad[pid] = f4(c.L, rad.w);
}

// Shades the queue; paths whose surface belongs to the second plugin group
// are moved to the side queue (their slots, count ctr[8]) for kShadeGroup2.
// Each block first sorts its paths by what they hit (miss, emitter, material
// type; counting sort in LDS), so that its waves run one plugin branch each
// instead of every branch their paths need.
// TABLE: the scene has at most kMatTableMax materials and the block keeps
// all of them in LDS (18 KiB); otherwise each thread stages its own hit's
// material (72 KiB per block, which halves the resident waves).
template <bool ONE, bool TABLE>
__global__ __launch_bounds__(kShadeBlock) void kShade(SceneView sc, BatchDesc b, Wavefront w, int cur) {
  __shared__ float4 sMat[TABLE ? kMatTableMax * 9 : kShadeBlock * 9];
  __shared__ int sSlot[kShadeBlock];
  __shared__ uint32_t sCnt[kShadeKeys];
  __shared__ BlockAppend sApp;
  const uint32_t count = w.ctr[cur];
  if (blockIdx.x * kShadeBlock >= count) return;  // whole block past the queue
  if (TABLE) {
    const float4* src = reinterpret_cast<const float4*>(sc.mats);
    for (uint32_t k = threadIdx.x; k < 9 * sc.numMats; k += kShadeBlock) sMat[k] = src[k];
  }
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < count;
  bool side = false;
  uint32_t key = kShadeKeys - 1;
  if (valid) {
    int2 hi = *reinterpret_cast<const int2*>(&w.hitI[i]);  // the hit carries its instance's shade word
    const uint32_t sw = uint32_t(hi.y);
    side = hi.x >= 0 && (sw & kShadeSide);
    key = hi.x < 0 ? 0u : (sw & kShadeLight) ? 1u : 2u + ((sw >> kShadeTypeShift) & kShadeTypeMask);
    if (side || key >= uint32_t(kShadeKeys - 1)) key = kShadeKeys - 1;
  }
  int si = waveAppend(&w.ctr[8], side);
  if (side) w.side[si] = int(i);
  if (threadIdx.x < kShadeKeys) sCnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t rank = key < uint32_t(kShadeKeys - 1) ? atomicAdd(&sCnt[key], 1u) : 0u;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int k = 0; k < kShadeKeys - 1; ++k) {
      uint32_t c = sCnt[k];
      sCnt[k] = run;
      run += c;
    }
    sCnt[kShadeKeys - 1] = run;
  }
  __syncthreads();
  if (key < uint32_t(kShadeKeys - 1)) sSlot[sCnt[key] + rank] = int(i);
  const uint32_t total = sCnt[kShadeKeys - 1];
  __syncthreads();
  const int mine = threadIdx.x < total ? sSlot[threadIdx.x] : -1;
  if (TABLE) shadeOne<ONE, false, kMatTable, true>(sc, b, w, cur, mine, sMat, &sApp);
  else shadeOne<ONE, false, kMatStaged, true>(sc, b, w, cur, mine, sMat + threadIdx.x * 9, &sApp);
}

// Second plugin group, grid-stride over the side queue.
template <bool ONE>
__global__ __launch_bounds__(256) void kShadeGroup2(SceneView sc, BatchDesc b, Wavefront w, int cur) {
  const uint32_t count = w.ctr[8];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
    shadeOne<ONE, true, kMatGlobal, false>(sc, b, w, cur, w.side[i], nullptr, nullptr);
}

// EXT: the extension fetch counters; SHADE: the counts shading appends to
// (next queue 1 - cur, shadow rays, side queue) and the shadow fetch counters.
template <bool EXT, bool SHADE>
__global__ void kResetCounters(Wavefront w, int cur) {
  if (threadIdx.x == 0) {
    if (SHADE) {
      w.ctr[1 - cur] = 0;
      w.ctr[2] = 0;
      w.ctr[4] = 0;
      w.ctr[8] = 0;
    }
    if (EXT) w.ctr[3] = 0;
  }
  if (threadIdx.x < 2 * kXcds && (threadIdx.x < kXcds ? EXT : SHADE)) w.ctr[kXcdCtr + threadIdx.x * 32] = 0;
}

// Ray counts of the bounce: every queued path traced one extension ray,
// every appended shadow ray is traced.
__global__ void kStatRays(Wavefront w, int cur) {
  if (threadIdx.x == 0) {
    w.cnt64[4] += w.ctr[cur];
    w.cnt64[5] += w.ctr[2];
  }
}

// Ordered fold of the batch's frames into the output images
// (raytrace.projective.rgen:146-178), one thread per pixel, one wave per
// block.  A pixel's frames are adjacent in the wavefront (pixel-major
// paths), so when the wave's 64 pixels belong to one work unit their frames
// are loaded in coalesced runs into LDS, kAccChunk frames at a time (rows
// padded by one entry against bank conflicts), and each lane folds its own
// row in frame order; otherwise each lane reads its frames directly.
// Chunks of 4 frames keep a block's LDS at 5 KB, so enough one-wave blocks
// are resident to hide each lane's serial fold: with 16 (17 KB, about two
// waves per SIMD) the 64-shot batches, 32 frames per pixel, spent 0.77 ms
// per accumulate of 2^25 paths, 14 % of the frame; 4 and 2 measure the same.
constexpr uint32_t kAccChunk = 4;
__device__ __forceinline__ void accFold(const WorkUnit& un, uint32_t f, float4 r, float4& mean, float& W) {
  float fw = r.w;
  f3 L = f3{clampf(r.x, 0.f, 10.f), clampf(r.y, 0.f, 10.f), clampf(r.z, 0.f, 10.f)};
  f3 rws = fw * L;
  if (un.f0 + int32_t(f) == 0) {
    f3 m = rws / fw;
    mean = f4(m, 1.f);
    W = fw;
  } else {
    f3 oldRws = xyz(mean) * W;
    float nW = W + fw;
    f3 m = (oldRws + rws) / nW;
    mean = f4(m, 1.f);
    W = nW;
  }
}
__global__ __launch_bounds__(64) void kAccumulate(BatchDesc b, Wavefront w) {
  __shared__ float4 sRad[64 * (kAccChunk + 1)];
  const uint32_t t = blockIdx.x * 64 + threadIdx.x, lane = threadIdx.x;
  const bool valid = t < b.nAcc;
  const uint32_t ui = valid ? findUnit<true>(b.units, b.nUnits, t) : 0xffffffffu;
  const bool together = __ballot(valid && ui == __shfl(ui, 0)) == ~0ull;  // wave-uniform
  if (!valid) return;  // only in the last wave, which then is not `together`
  const WorkUnit un = b.units[ui];
  const uint32_t k = t - un.accBase;
  const ShotParams& sp = b.shotTab[un.shot];
  const uint32_t pix = runPixel(un.pix0 + k, sp.width, sp.height);
  size_t plane = size_t(sp.width) * sp.height;
  float4* img0 = sp.images;
  float4* img8 = sp.images + 8 * plane;
  float4 mean = img0[pix];
  float W = img8[pix].x;
  if (together) {
    const uint32_t k0 = k - lane;  // the wave's first pixel of the unit
    const float4* src = w.rad + un.base + size_t(k0) * un.nf;
    for (uint32_t c0 = 0; c0 < un.nf; c0 += kAccChunk) {
      const uint32_t cnt = min(kAccChunk, un.nf - c0);
      __syncthreads();
      for (uint32_t e = lane; e < 64 * cnt; e += 64) {
        const uint32_t px = e / cnt, fr = e - px * cnt;
        sRad[px * (cnt + 1) + fr] = src[size_t(px) * un.nf + c0 + fr];
      }
      __syncthreads();
      for (uint32_t fr = 0; fr < cnt; ++fr) accFold(un, c0 + fr, sRad[lane * (cnt + 1) + fr], mean, W);
    }
  } else {
    for (uint32_t f = 0; f < un.nf; ++f) accFold(un, f, w.rad[un.base + k * un.nf + f], mean, W);
  }
  img0[pix] = mean;
  img8[pix] = make_float4(W, W, W, W);
}

__global__ __launch_bounds__(256) void kFirstHitState(SceneView sc, const ShotParams* sp, Wavefront w, ash_hit* out,
                                                      uint32_t P) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  int4 hi = w.hitI[p];  // slot p of queue 0 after raygen is path p
  ash_hit o;
  if (hi.x < 0) {
    o.t = kInf;
    o.u = o.v = 0.f;
    o.inst_id = -1;
    o.prim_id = -1;
    for (int k = 0; k < 3; ++k) o.normal[k] = 0.f, o.pos[k] = 0.f;
  } else {
    float4 hf = w.hitF[p];
    Hit h;
    hitState(sc, hi.x, uint32_t(hi.y), hi.w, hf.y, hf.z, xyz(w.rayD[0][p]), 0u, h);
    o.t = hf.x;
    o.u = hf.y;
    o.v = hf.z;
    o.inst_id = hi.x;
    o.prim_id = hi.z - int(sc.inst[hi.x].triBase);  // global id minus the instance's first triangle
    o.normal[0] = h.N.x, o.normal[1] = h.N.y, o.normal[2] = h.N.z;
    o.pos[0] = h.pos.x, o.pos[1] = h.pos.y, o.pos[2] = h.pos.z;
  }
  out[runPixel(p, sp->width, sp->height)] = o;
}

}  // namespace

void launchRaygen(const BatchDesc& b, const Wavefront& w, hipStream_t s) {
  kRaygen<<<(b.n + 255) / 256, 256, 0, s>>>(b, w);
}
template <bool COUNT, int TREE>
static void extendWave(const SceneView& sc, const Wavefront& w, int cur, bool shallow, bool xcd, uint32_t grid,
                       hipStream_t s) {
  if (xcd) {
    if (shallow) kExtend<COUNT, TREE, true, kStackLdsMin><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    else kExtend<COUNT, TREE, true, kStackLds><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
  } else {
    if (shallow) kExtend<COUNT, TREE, false, kStackLdsMin><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    else kExtend<COUNT, TREE, false, kStackLds><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
  }
}
// Closest-hit kernel per tree and fetch scheme: per-XCD slices walk float
// 4-wide nodes with dynamic fetch (kExtendDyn), quantized nodes with wave
// fetch; queue order uses wave fetch; the per-lane refill scheme walks the
// 128-byte nodes (kept beside the quantized ones).
constexpr int kDynRefill = 48;
template <bool COUNT>
static void extendFor(const SceneView& sc, const Wavefront& w, int cur, bool shallow, uint32_t scheme, uint32_t grid,
                      hipStream_t s) {
  const bool wide = sc.nodes4 != nullptr, xcd = scheme == kSchemeXcdSlices;
  if (!wide) {
    if (xcd) kExtend<COUNT, 0, true, kStackLds><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    else kExtend<COUNT, 0, false, kStackLds><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
  } else if (scheme == kSchemeRefill) {
    if (shallow) kExtendWide<COUNT, kStackLdsMin><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    else kExtendWide<COUNT, kStackLds><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
  } else if (sc.nodesQ) {
    extendWave<COUNT, 2>(sc, w, cur, shallow, xcd, grid, s);
  } else if (xcd) {
    if (shallow) kExtendDyn<COUNT, kStackLdsMin, kDynRefill><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    else kExtendDyn<COUNT, kStackLds, kDynRefill><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
  } else {
    extendWave<COUNT, 1>(sc, w, cur, shallow, false, grid, s);
  }
}
void launchExtend(const SceneView& sc, const Wavefront& w, int cur, uint32_t ldsStack, uint32_t scheme,
                  uint32_t grid, bool count, hipStream_t s) {
  if (sc.twoLevel) {
    if (count) kExtend2L<true><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    else kExtend2L<false><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    return;
  }
  const bool shallow = ldsStack == uint32_t(kStackLdsMin);
  if (count) extendFor<true>(sc, w, cur, shallow, scheme, grid, s);
  else extendFor<false>(sc, w, cur, shallow, scheme, grid, s);
}
bool canPacketCamera(const SceneView& sc, uint32_t scheme) {
  return !sc.twoLevel && scheme == kSchemeXcdSlices && sc.nodes4 != nullptr && sc.nodesQ == nullptr;
}
void launchExtendCamera(const SceneView& sc, const Wavefront& w, uint32_t ldsStack, uint32_t scheme, uint32_t grid,
                        bool count, bool coarse, hipStream_t s) {
  if (!canPacketCamera(sc, scheme)) {
    launchExtend(sc, w, 0, ldsStack, scheme, grid, count, s);
    return;
  }
  if (coarse) {
    if (count) kExtendPacket<true, 256><<<grid, kTraceBlock, 0, s>>>(sc, w, 0);
    else kExtendPacket<false, 256><<<grid, kTraceBlock, 0, s>>>(sc, w, 0);
  } else {
    if (count) kExtendPacket<true, 64><<<grid, kTraceBlock, 0, s>>>(sc, w, 0);
    else kExtendPacket<false, 64><<<grid, kTraceBlock, 0, s>>>(sc, w, 0);
  }
}
void launchShade(const SceneView& sc, const BatchDesc& b, const Wavefront& w, int cur, uint32_t maxRays,
                 hipStream_t s) {
  const uint32_t blocks = (maxRays + kShadeBlock - 1) / kShadeBlock;
  const bool table = sc.numMats <= kMatTableMax;
  if (b.shots == 1) {
    if (table) kShade<true, true><<<blocks, kShadeBlock, 0, s>>>(sc, b, w, cur);
    else kShade<true, false><<<blocks, kShadeBlock, 0, s>>>(sc, b, w, cur);
    kShadeGroup2<true><<<512, 256, 0, s>>>(sc, b, w, cur);
  } else {
    if (table) kShade<false, true><<<blocks, kShadeBlock, 0, s>>>(sc, b, w, cur);
    else kShade<false, false><<<blocks, kShadeBlock, 0, s>>>(sc, b, w, cur);
    kShadeGroup2<false><<<512, 256, 0, s>>>(sc, b, w, cur);
  }
  kStatRays<<<1, 64, 0, s>>>(w, cur);
}
template <bool COUNT>
static void shadowFor(const SceneView& sc, const Wavefront& w, uint32_t scheme, uint32_t grid, hipStream_t s) {
  const bool wide = sc.nodes4 != nullptr, quant = sc.nodesQ != nullptr, xcd = scheme == kSchemeXcdSlices;
  if (wide && scheme == kSchemeRefill) kShadowWide<COUNT><<<grid, kTraceBlock, 0, s>>>(sc, w);
  else if (quant && xcd) kShadow<COUNT, 2, true><<<grid, kTraceBlock, 0, s>>>(sc, w);
  else if (quant) kShadow<COUNT, 2, false><<<grid, kTraceBlock, 0, s>>>(sc, w);
  else if (wide && xcd) kShadow<COUNT, 1, true><<<grid, kTraceBlock, 0, s>>>(sc, w);
  else if (wide) kShadow<COUNT, 1, false><<<grid, kTraceBlock, 0, s>>>(sc, w);
  else if (xcd) kShadow<COUNT, 0, true><<<grid, kTraceBlock, 0, s>>>(sc, w);
  else kShadow<COUNT, 0, false><<<grid, kTraceBlock, 0, s>>>(sc, w);
}
void launchShadow(const SceneView& sc, const Wavefront& w, uint32_t scheme, uint32_t grid, bool count,
                  hipStream_t s) {
  if (sc.twoLevel) {
    if (count) kShadow2L<true><<<grid, kTraceBlock, 0, s>>>(sc, w);
    else kShadow2L<false><<<grid, kTraceBlock, 0, s>>>(sc, w);
    return;
  }
  if (count) shadowFor<true>(sc, w, scheme, grid, s);
  else shadowFor<false>(sc, w, scheme, grid, s);
}
void launchAccumulate(const BatchDesc& b, const Wavefront& w, hipStream_t s) {
  kAccumulate<<<(b.nAcc + 63) / 64, 64, 0, s>>>(b, w);
}
void launchSkyPrepare(DevSky* sky, hipStream_t s) { kSkyPrepare<<<1, 64, 0, s>>>(sky); }

void launchResetCounters(const Wavefront& w, int cur, hipStream_t s) { kResetCounters<true, true><<<1, 64, 0, s>>>(w, cur); }
void launchResetShade(const Wavefront& w, int cur, hipStream_t s) { kResetCounters<false, true><<<1, 64, 0, s>>>(w, cur); }
void launchResetExtend(const Wavefront& w, hipStream_t s) { kResetCounters<true, false><<<1, 64, 0, s>>>(w, 0); }
bool canFuseTrace(const SceneView& sc, uint32_t scheme) {
  return sc.twoLevel || (scheme == kSchemeXcdSlices && sc.nodes4 != nullptr && sc.nodesQ == nullptr);
}
void launchTrace(const SceneView& sc, const Wavefront& w, int cur, uint32_t ldsStack, uint32_t grid, bool count,
                 hipStream_t s) {
  if (sc.twoLevel) {
    if (count) kTrace2L<true><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    else kTrace2L<false><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    return;
  }
  const bool shallow = ldsStack == uint32_t(kStackLdsMin);
  if (count) {
    if (shallow) kTraceDyn<true, kStackLdsMin, kDynRefill><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    else kTraceDyn<true, kStackLds, kDynRefill><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
  } else {
    if (shallow) kTraceDyn<false, kStackLdsMin, kDynRefill><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
    else kTraceDyn<false, kStackLds, kDynRefill><<<grid, kTraceBlock, 0, s>>>(sc, w, cur);
  }
}
void launchFirstHit(const SceneView& sc, const ShotParams* shot, const Wavefront& w, ash_hit* out, uint32_t P,
                    hipStream_t s) {
  kFirstHitState<<<(P + 255) / 256, 256, 0, s>>>(sc, shot, w, out, P);
}

}  // namespace ashd
