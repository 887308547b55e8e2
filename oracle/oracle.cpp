// oracle.cpp -- CPU restatement of Asuna's render path (TEST INFRASTRUCTURE).
//
// This file is the checker the HIP product is compared against.  It is a
// plain, per-path restatement of the GLSL programs in src/shaders (one
// function per shader function, in the shader's own evaluation order), so
// that a reader can check it line by line against the reference.  It is
// compiled with -ffp-contract=off and uses only IEEE basic operations where
// the reference does, so that the FP32 instantiation reproduces every
// rounding of the same operation sequence.
//
// Every part is pinned by a closed form, a known answer or an invariant in
// tests/test_oracle.py, among them the argument permutation of the
// rough-plastic NEE (brdf_rough_plastic.rchit:234-240: eval at V = L = N
// with distinct ax, ay and 1/ior^2) and kang18's object-space normal and
// tangent maps (brdf_kang18.rchit: maps encoding the plane's own object
// frame give the world frame of the transformed geometry).
//
// Nothing in the product links or calls this file.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

// ------------------------------------------------------------------ vec3
template <class R>
struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  explicit V3(R a) : x(a), y(a), z(a) {}
  R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class R> R length(V3<R> a) { return std::sqrt(dot(a, a)); }
// GLSL normalize: v * (1 / length(v)).  Both implementations of this repo use
// this exact formula (the vendor's is not documented).
template <class R> V3<R> normalize(V3<R> a) {
  R inv = R(1) / std::sqrt(dot(a, a));
  return a * inv;
}
template <class R> V3<R> vabs(V3<R> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class R> V3<R> vmin(V3<R> a, V3<R> b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
template <class R> V3<R> vmax(V3<R> a, V3<R> b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
template <class R> R clampr(R x, R lo, R hi) { return std::min(std::max(x, lo), hi); }
template <class R> R sqr(R x) { return x * x; }
template <class R> V3<R> load3(const float* p) { return {R(p[0]), R(p[1]), R(p[2])}; }

// math.glsl:4-15 constants are float literals in GLSL.
template <class R> struct K {
  static constexpr R PI = R(3.14159265358979323846);
  static constexpr R TWO_PI = R(6.28318530717958647692);
  static constexpr R INV_PI = R(0.31830988618379067154);
  static constexpr R INV_2PI = R(0.15915494309189533577);
  static constexpr R INV_4PI = R(0.07957747154594766788);
  static constexpr R PI_OVER_2 = R(1.57079632679489661923);
  static constexpr R PI_OVER_4 = R(0.78539816339744830961);
  static constexpr R EPS = R(0.001);
  static constexpr R INF = R(10000000000.0);
  static constexpr R MINIMUM = R(0.00001);
};

// structs.glsl:16-46 flag sets
constexpr uint32_t EBsdfNull = 0, EDiffuseReflection = 1, EDiffuseTransmission = 2,
                   EGlossyReflection = 4, EGlossyTransmission = 8, ESpecularReflection = 16,
                   ESpecularTransmission = 32;
constexpr uint32_t ESmooth = EDiffuseReflection | EDiffuseTransmission | EGlossyReflection |
                             EGlossyTransmission;
constexpr uint32_t EDelta = 1, EArea = 2;
inline bool isNonSpecular(uint32_t f) { return (f & ESmooth) != 0; }

// ------------------------------------------------------------------- RNG
// math.glsl:20-30
uint32_t xxhash32Seed(uint32_t px, uint32_t py, uint32_t pz) {
  const uint32_t p0 = 2246822519U, p1 = 3266489917U, p2 = 668265263U, p3 = 374761393U;
  uint32_t h32 = pz + p3 + px * p1;
  h32 = p2 * ((h32 << 17) | (h32 >> (32 - 17)));
  h32 += py * p1;
  h32 = p2 * ((h32 << 17) | (h32 >> (32 - 17)));
  h32 = p0 * (h32 ^ (h32 >> 15));
  h32 = p1 * (h32 ^ (h32 >> 13));
  return h32 ^ (h32 >> 16);
}
// math.glsl:34-39
uint32_t pcg(uint32_t& state) {
  uint32_t prev = state * 747796405u + 2891336453u;
  uint32_t word = ((prev >> ((prev >> 28u) + 4u)) ^ prev) * 277803737u;
  state = prev;
  return (word >> 22u) ^ word;
}
// math.glsl:42: float(pcg) * (1.0 / float(0xffffffffu)); float(0xffffffff)
// rounds to 2^32, so the FP32 draw is float(u) * 2^-32.  The FP64 estimator
// draws u / (2^32 - 1).
template <class R> R rnd(uint32_t& s);
template <> float rnd<float>(uint32_t& s) { return float(pcg(s)) * (1.0f / 4294967296.0f); }
template <> double rnd<double>(uint32_t& s) { return double(pcg(s)) * (1.0 / 4294967295.0); }

// ---------------------------------------------------------------- matrix
// Column-major 4x4 as GLSL/nvmath store it: m[c*4 + r].
template <class R> V3<R> transformPoint(const float* m, V3<R> p) {
  // math.glsl:46-50 (transform * vec4(p,1)).xyz / w
  R x = R(m[0]) * p.x + R(m[4]) * p.y + R(m[8]) * p.z + R(m[12]);
  R y = R(m[1]) * p.x + R(m[5]) * p.y + R(m[9]) * p.z + R(m[13]);
  R z = R(m[2]) * p.x + R(m[6]) * p.y + R(m[10]) * p.z + R(m[14]);
  R w = R(m[3]) * p.x + R(m[7]) * p.y + R(m[11]) * p.z + R(m[15]);
  return V3<R>(x, y, z) / w;
}
template <class R> V3<R> transformVector(const float* m, V3<R> v) {
  // math.glsl:52-56 (w = 0)
  R x = R(m[0]) * v.x + R(m[4]) * v.y + R(m[8]) * v.z;
  R y = R(m[1]) * v.x + R(m[5]) * v.y + R(m[9]) * v.z;
  R z = R(m[2]) * v.x + R(m[6]) * v.y + R(m[10]) * v.z;
  return {x, y, z};
}
template <class R> V3<R> makeNormal(V3<R> n) {  // math.glsl:58-61
  if (length(n) == R(0)) return n;
  return normalize(n);
}
template <class R> V3<R> transformDirection(const float* m, V3<R> d) {  // math.glsl:63-66
  return makeNormal(transformVector(m, d));
}

// ------------------------------------------------------------ warps, misc
template <class R> R safeSqrt(R v) { return std::sqrt(std::max(R(0), v)); }
template <class R> V3<R> toWorld(V3<R> X, V3<R> Y, V3<R> Z, V3<R> v) {  // math.glsl:72-74
  return v.x * X + v.y * Y + v.z * Z;
}
template <class R> V3<R> toLocal(V3<R> X, V3<R> Y, V3<R> Z, V3<R> v) {  // math.glsl:76-78
  return {dot(v, X), dot(v, Y), dot(v, Z)};
}
template <class R> V3<R> uniformSampleSphere(R u0, R u1) {  // math.glsl:133-138
  R z = R(1) - R(2) * u0;
  R r = std::sqrt(std::max(R(0), R(1) - z * z));
  R phi = K<R>::TWO_PI * u1;
  return {r * std::cos(phi), r * std::sin(phi), z};
}
template <class R> void concentricSampleDisk(R u0, R u1, R& ox, R& oy) {  // math.glsl:148-163
  R ux = R(2) * u0 - R(1), uy = R(2) * u1 - R(1);
  if (ux == R(0) && uy == R(0)) { ox = 0; oy = 0; return; }
  R theta, r;
  if (std::fabs(ux) > std::fabs(uy)) {
    r = ux;
    theta = K<R>::PI_OVER_4 * (uy / ux);
  } else {
    r = uy;
    theta = K<R>::PI_OVER_2 - K<R>::PI_OVER_4 * (ux / uy);
  }
  ox = r * std::cos(theta);
  oy = r * std::sin(theta);
}
template <class R> V3<R> cosineSampleHemisphere(R u0, R u1) {  // math.glsl:165-169
  R dx, dy;
  concentricSampleDisk(u0, u1, dx, dy);
  R z = std::sqrt(std::max(R(0), R(1) - dx * dx - dy * dy));
  return {dx, dy, z};
}
template <class R> R cosineHemispherePdf(R c) {  // math.glsl:171-174
  if (c <= R(0)) return 0;
  return c * K<R>::INV_PI;
}
template <class R> R powerHeuristic(R a, R b) {  // math.glsl:184-189
  a = a * a;
  b = b * b + a;
  if (b == R(0)) return 0;
  return a / b;
}
template <class R> void basis(V3<R> n, V3<R>& f, V3<R>& r) {  // math.glsl:199-216
  if (n.z < R(-0.999999)) {
    f = V3<R>(0, -1, 0);
    r = V3<R>(-1, 0, 0);
  } else {
    R a = R(1) / (R(1) + n.z);
    R b = -n.x * n.y * a;
    f = V3<R>(R(1) - n.x * n.x * a, b, -n.x);
    r = V3<R>(b, R(1) - n.y * n.y * a, -n.y);
  }
}
// math.glsl:241-266 (Waechter & Binder, Ray Tracing Gems ch. 6).  The
// integer-ulp offset is defined on float bits; the FP64 estimator applies it
// to the float-rounded position.
inline float offsetComp(float p, float n) {
  int32_t of = int32_t(256.0f * n);
  int32_t bits;
  std::memcpy(&bits, &p, 4);
  bits += (p < 0) ? -of : of;
  float pi;
  std::memcpy(&pi, &bits, 4);
  return std::fabs(p) < (1.0f / 32.0f) ? p + (1.0f / 65536.0f) * n : pi;
}
template <class R> V3<R> offsetPositionAlongNormal(V3<R> p, V3<R> n) {
  return V3<R>(R(offsetComp(float(p.x), float(n.x))), R(offsetComp(float(p.y), float(n.y))),
               R(offsetComp(float(p.z), float(n.z))));
}
template <class R> V3<R> reflect(V3<R> i, V3<R> n) { return i - R(2) * dot(n, i) * n; }
template <class R> V3<R> refract(V3<R> i, V3<R> n, R eta) {
  R d = dot(n, i);
  R k = R(1) - eta * eta * (R(1) - d * d);
  if (k < R(0)) return V3<R>(0);
  return eta * i - (eta * d + std::sqrt(k)) * n;
}

// ----------------------------------------------------- texture emulation
// texture(sampler2D, uv) at LOD 0 with VK_FILTER_LINEAR and the zero-init
// address mode (VK_SAMPLER_ADDRESS_MODE_REPEAT), texture.cpp:244-248;
// Vulkan spec "Texel Filtering": u' = u*W - 0.5, i0 = floor(u'), alpha =
// frac(u'), wrap i mod W.
struct Img {
  uint32_t w = 1, h = 1;
  std::vector<float> px = std::vector<float>(4, 0.f);  // rgba
  // Marginal table: only column 0 holds data (texture.cpp:209-214).
  bool col0_only = false;
};
inline int wrapi(int i, int n) {
  int m = i % n;
  return m < 0 ? m + n : m;
}
template <class R> void texel(const Img& im, int x, int y, R out[4]) {
  if (im.col0_only && x != 0) {
    out[0] = out[1] = out[2] = out[3] = 0;
    return;
  }
  const float* p = &im.px[(size_t(y) * (im.col0_only ? 1 : im.w) + (im.col0_only ? 0 : x)) * 4];
  for (int c = 0; c < 4; ++c) out[c] = R(p[c]);
}
template <class R> void tex2D(const Img& im, R u, R v, R out[4]) {
  if (!(u == u)) u = 0;  // NaN coordinates fetch texel (0,0)-ish instead of UB
  if (!(v == v)) v = 0;
  R x = u * R(im.w) - R(0.5), y = v * R(im.h) - R(0.5);
  R fx = std::floor(x), fy = std::floor(y);
  R a = x - fx, b = y - fy;
  int x0 = wrapi(int(fx), int(im.w)), y0 = wrapi(int(fy), int(im.h));
  int x1 = wrapi(int(fx) + 1, int(im.w)), y1 = wrapi(int(fy) + 1, int(im.h));
  R t00[4], t10[4], t01[4], t11[4];
  texel(im, x0, y0, t00);
  texel(im, x1, y0, t10);
  texel(im, x0, y1, t01);
  texel(im, x1, y1, t11);
  for (int c = 0; c < 4; ++c)
    out[c] = (R(1) - a) * (R(1) - b) * t00[c] + a * (R(1) - b) * t10[c] +
             (R(1) - a) * b * t01[c] + a * b * t11[c];
}

// ------------------------------------------------------- env tables
// EnvMap::EnvMap, texture.cpp:144-226, with its mixed float/double
// arithmetic (0.3 etc. and 1e-7 are double literals).
void envTables(uint32_t width, uint32_t height, const float* pixel, std::vector<float>& marg,
               std::vector<float>& cond) {
  int W = int(width), H = int(height);
  std::vector<float> pdf2D(size_t(W) * H), cdf2D(size_t(W) * H), pdf1D(H), cdf1D(H);
  marg.assign(size_t(H) * 2, 0.f);
  cond.assign(size_t(W) * H * 2, 0.f);
  float weightSum = 0.0f;
  for (int j = 0; j < H; j++) {
    float rowWeightSum = 0.0f;
    for (int i = 0; i < W; i++) {
      const float* c = &pixel[(size_t(j) * W + i) * 4];
      float weight = float(0.3 * c[0] + 0.6 * c[1] + 0.1 * c[2]);
      rowWeightSum += weight;
      pdf2D[size_t(j) * W + i] = weight;
      cdf2D[size_t(j) * W + i] = rowWeightSum;
    }
    for (int i = 0; i < W; i++) {
      pdf2D[size_t(j) * W + i] = float(double(pdf2D[size_t(j) * W + i]) / (double(rowWeightSum) + 1e-7));
      cdf2D[size_t(j) * W + i] = float(double(cdf2D[size_t(j) * W + i]) / (double(rowWeightSum) + 1e-7));
    }
    weightSum += rowWeightSum;
    pdf1D[j] = rowWeightSum;
    cdf1D[j] = weightSum;
  }
  for (int j = 0; j < H; j++) {
    cdf1D[j] = float(double(cdf1D[j]) / (double(weightSum) + 1e-7));
    pdf1D[j] = float(double(pdf1D[j]) / (double(weightSum) + 1e-7));
  }
  auto lowerBound = [](const float* a, int lo, int hi, float value) {
    while (lo < hi) {
      int mid = (lo + hi) / 2;
      if (a[mid] < value) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  for (int j = 0; j < H; j++) {
    float invHeight = float(j + 1) / float(H);
    float row = float(lowerBound(cdf1D.data(), 0, H, invHeight));
    marg[size_t(j) * 2 + 0] = row / float(H);
    marg[size_t(j) * 2 + 1] = pdf1D[j];
  }
  for (int j = 0; j < H; j++) {
    for (int i = 0; i < W; i++) {
      float invWidth = float(i + 1) / float(W);
      float col = float(lowerBound(cdf2D.data(), j * W, (j + 1) * W, invWidth) - j * W);
      cond[(size_t(j) * W + i) * 2 + 0] = col / float(W);
      cond[(size_t(j) * W + i) * 2 + 1] = pdf2D[size_t(j) * W + i];
    }
  }
}

}  // namespace

// ================================================================= scene
struct orc_scene {
  struct Inst {
    float o2w[12];  // row-major 3x4 (gl_ObjectToWorldEXT)
    float w2o[12];  // row-major 3x4 (gl_WorldToObjectEXT)
    uint32_t mesh, material;
    int32_t light;
  };
  std::vector<Inst> inst;
  std::vector<std::vector<GpuVertex>> verts;
  std::vector<std::vector<uint32_t>> idx;
  std::vector<GpuMaterial> mats;
  std::vector<GpuLight> lights;
  std::vector<Img> textures;
  bool hasEnv = false;
  Img env, marg, cond;
  GpuSunAndSky sky{};  // in_use 0 until orc_set_sun_and_sky
  // Flattened world-space triangles (9 floats) with (instance, primitive).
  std::vector<float> tri;
  std::vector<int32_t> triInst, triPrim;
  // Binned-SAH BVH (the oracle's own, unrelated to the product's builder).
  struct Node {
    float lo[3], hi[3];
    uint32_t left;   // internal: first child index (children are left, left+1); leaf: first prim
    uint32_t count;  // 0 = internal
  };
  std::vector<Node> nodes;
  std::vector<uint32_t> order;  // leaf prim order -> global triangle
};

namespace {

void invert3x4(const float* m /*row-major 3x4*/, float* out) {
  // Affine inverse in double, rounded once (the driver's gl_WorldToObjectEXT
  // rounding is not documented).
  double a[3][3], t[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) a[r][c] = m[r * 4 + c];
    t[r] = m[r * 4 + 3];
  }
  double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
               a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
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

// Object->world point with the fixed operation order this repo uses for
// world-space triangles: ((m0*x + m1*y) + m2*z) + m3 per row.
inline void xformPoint3x4(const float* m, const float* p, float* o) {
  for (int r = 0; r < 3; ++r)
    o[r] = m[r * 4 + 0] * p[0] + m[r * 4 + 1] * p[1] + m[r * 4 + 2] * p[2] + m[r * 4 + 3];
}

// ----------------------------------------------------- BVH (binned SAH)
struct Build {
  orc_scene* s;
  std::vector<float> cen;  // centroids
  std::vector<float> blo, bhi;
};

void buildNode(Build& b, uint32_t ni, uint32_t first, uint32_t count, int depth) {
  orc_scene& s = *b.s;
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
  for (uint32_t i = first; i < first + count; ++i) {
    uint32_t t = s.order[i];
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], b.blo[t * 3 + k]);
      hi[k] = std::max(hi[k], b.bhi[t * 3 + k]);
      clo[k] = std::min(clo[k], b.cen[t * 3 + k]);
      chi[k] = std::max(chi[k], b.cen[t * 3 + k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    s.nodes[ni].lo[k] = lo[k];
    s.nodes[ni].hi[k] = hi[k];
  }
  auto area = [](const float* l, const float* h) {
    float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
    if (dx < 0) return 0.f;
    return 2.f * (dx * dy + dy * dz + dz * dx);
  };
  if (count <= 4 || depth > 60) {
    s.nodes[ni].left = first;
    s.nodes[ni].count = count;
    return;
  }
  const int NB = 16;
  float bestCost = 1e30f;
  int bestAxis = -1, bestSplit = -1;
  for (int ax = 0; ax < 3; ++ax) {
    float ext = chi[ax] - clo[ax];
    if (!(ext > 0)) continue;
    float bl[NB][3], bh[NB][3];
    uint32_t bc[NB] = {0};
    for (int i = 0; i < NB; ++i)
      for (int k = 0; k < 3; ++k) bl[i][k] = 1e30f, bh[i][k] = -1e30f;
    for (uint32_t i = first; i < first + count; ++i) {
      uint32_t t = s.order[i];
      int bi = std::min(NB - 1, int((b.cen[t * 3 + ax] - clo[ax]) / ext * NB));
      bc[bi]++;
      for (int k = 0; k < 3; ++k) {
        bl[bi][k] = std::min(bl[bi][k], b.blo[t * 3 + k]);
        bh[bi][k] = std::max(bh[bi][k], b.bhi[t * 3 + k]);
      }
    }
    float rl[NB][3], rh[NB][3];
    uint32_t rc[NB];
    float cl[3] = {1e30f, 1e30f, 1e30f}, ch[3] = {-1e30f, -1e30f, -1e30f};
    uint32_t cc = 0;
    for (int i = NB - 1; i > 0; --i) {
      cc += bc[i];
      for (int k = 0; k < 3; ++k) cl[k] = std::min(cl[k], bl[i][k]), ch[k] = std::max(ch[k], bh[i][k]);
      rc[i] = cc;
      for (int k = 0; k < 3; ++k) rl[i][k] = cl[k], rh[i][k] = ch[k];
    }
    float ll[3] = {1e30f, 1e30f, 1e30f}, lh[3] = {-1e30f, -1e30f, -1e30f};
    uint32_t lc = 0;
    for (int i = 0; i < NB - 1; ++i) {
      lc += bc[i];
      for (int k = 0; k < 3; ++k) ll[k] = std::min(ll[k], bl[i][k]), lh[k] = std::max(lh[k], bh[i][k]);
      if (lc == 0 || rc[i + 1] == 0) continue;
      float cost = area(ll, lh) * lc + area(rl[i + 1], rh[i + 1]) * rc[i + 1];
      if (cost < bestCost) bestCost = cost, bestAxis = ax, bestSplit = i;
    }
  }
  uint32_t mid;
  if (bestAxis < 0) {
    mid = first + count / 2;
  } else {
    float ext = chi[bestAxis] - clo[bestAxis];
    auto it = std::partition(s.order.begin() + first, s.order.begin() + first + count,
                             [&](uint32_t t) {
                               int bi = std::min(NB - 1, int((b.cen[t * 3 + bestAxis] - clo[bestAxis]) / ext * NB));
                               return bi <= bestSplit;
                             });
    mid = uint32_t(it - s.order.begin());
    if (mid == first || mid == first + count) mid = first + count / 2;
  }
  uint32_t l = uint32_t(s.nodes.size());
  s.nodes.push_back({});
  s.nodes.push_back({});
  s.nodes[ni].left = l;
  s.nodes[ni].count = 0;
  buildNode(b, l, first, mid - first, depth + 1);
  buildNode(b, l + 1, mid, first + count - mid, depth + 1);
}

// ------------------------------------------------ watertight intersection
// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1) 2013.
// Barycentrics: U,V,W weight v0,v1,v2; GLSL _bary = (V, W)/det.
template <class R> struct RayW {
  V3<R> o, d;
  int kx, ky, kz;
  R Sx, Sy, Sz;
  void setup() {
    R ad[3] = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
    kz = (ad[0] > ad[1]) ? (ad[0] > ad[2] ? 0 : 2) : (ad[1] > ad[2] ? 1 : 2);
    kx = kz + 1 == 3 ? 0 : kz + 1;
    ky = kx + 1 == 3 ? 0 : kx + 1;
    if (d[kz] < R(0)) std::swap(kx, ky);
    Sx = d[kx] / d[kz];
    Sy = d[ky] / d[kz];
    Sz = R(1) / d[kz];
  }
};

template <class R>
bool intersectTri(const RayW<R>& r, const float* tv, R tmin, R tmax, R& t, R& bu, R& bv) {
  V3<R> A = V3<R>(R(tv[0]), R(tv[1]), R(tv[2])) - r.o;
  V3<R> B = V3<R>(R(tv[3]), R(tv[4]), R(tv[5])) - r.o;
  V3<R> C = V3<R>(R(tv[6]), R(tv[7]), R(tv[8])) - r.o;
  R Ax = A[r.kx] - r.Sx * A[r.kz], Ay = A[r.ky] - r.Sy * A[r.kz];
  R Bx = B[r.kx] - r.Sx * B[r.kz], By = B[r.ky] - r.Sy * B[r.kz];
  R Cx = C[r.kx] - r.Sx * C[r.kz], Cy = C[r.ky] - r.Sy * C[r.kz];
  R U = Cx * By - Cy * Bx;
  R V = Ax * Cy - Ay * Cx;
  R W = Bx * Ay - By * Ax;
  if (sizeof(R) == 4 && (U == R(0) || V == R(0) || W == R(0))) {
    U = R(double(Cx) * double(By) - double(Cy) * double(Bx));
    V = R(double(Ax) * double(Cy) - double(Ay) * double(Cx));
    W = R(double(Bx) * double(Ay) - double(By) * double(Ax));
  }
  if ((U < R(0) || V < R(0) || W < R(0)) && (U > R(0) || V > R(0) || W > R(0))) return false;
  R det = U + V + W;
  if (det == R(0)) return false;
  R Az = r.Sz * A[r.kz], Bz = r.Sz * B[r.kz], Cz = r.Sz * C[r.kz];
  R T = U * Az + V * Bz + W * Cz;
  R rcp = R(1) / det;
  R tt = T * rcp;
  if (!(tt >= tmin && tt <= tmax)) return false;
  t = tt;
  bu = V * rcp;
  bv = W * rcp;
  return true;
}

// Conservative slab test for the oracle's own BVH (boxes padded so that no
// hit the brute-force loop finds is culled).
template <class R>
bool hitBox(const orc_scene::Node& n, const V3<R>& o, const V3<R>& inv, R tmin, R tmax) {
  R t0 = tmin, t1 = tmax;
  for (int k = 0; k < 3; ++k) {
    R pad = (R(n.hi[k]) - R(n.lo[k])) * R(1e-5) + R(1e-6) * (std::fabs(R(n.hi[k])) + std::fabs(R(n.lo[k])) + R(1));
    R a = (R(n.lo[k]) - pad - o[k]) * inv[k];
    R b = (R(n.hi[k]) + pad - o[k]) * inv[k];
    if (std::isnan(a) || std::isnan(b)) {  // 0 * inf: ray parallel and on the slab plane
      if (o[k] < R(n.lo[k]) - pad || o[k] > R(n.hi[k]) + pad) return false;
      continue;
    }
    if (a > b) std::swap(a, b);
    t0 = std::max(t0, a);
    t1 = std::min(t1, b * R(1.0000005));
    if (t0 > t1) return false;
  }
  return true;
}

struct HitR {
  bool hit = false;
  double t = 0;
  double u = 0, v = 0;
  int32_t tri = -1;
};

template <class R>
HitR traceClosest(const orc_scene& s, V3<R> o, V3<R> d, R tmin, R tmax, bool brute) {
  RayW<R> r;
  r.o = o;
  r.d = d;
  r.setup();
  HitR best;
  R bt = tmax, bu = 0, bv = 0;
  int32_t bid = -1;
  auto testTri = [&](uint32_t g) {
    R t, u, v;
    if (intersectTri(r, &s.tri[size_t(g) * 9], tmin, tmax, t, u, v)) {
      // Closest hit; equal t resolved by the smaller global triangle id so the
      // result is independent of traversal order.
      if (bid < 0 || t < bt || (t == bt && int32_t(g) < bid)) {
        bt = t;
        bu = u;
        bv = v;
        bid = int32_t(g);
      }
    }
  };
  if (brute || s.nodes.empty()) {
    for (uint32_t g = 0; g < s.triInst.size(); ++g) testTri(g);
  } else {
    V3<R> inv(R(1) / d.x, R(1) / d.y, R(1) / d.z);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const orc_scene::Node& n = s.nodes[stack[--sp]];
      if (!hitBox(n, o, inv, tmin, bid >= 0 ? bt : tmax)) continue;
      if (n.count) {
        for (uint32_t i = 0; i < n.count; ++i) testTri(s.order[n.left + i]);
      } else {
        stack[sp++] = n.left;
        stack[sp++] = n.left + 1;
      }
    }
  }
  if (bid >= 0) {
    best.hit = true;
    best.t = double(bt);
    best.u = double(bu);
    best.v = double(bv);
    best.tri = bid;
  }
  return best;
}

template <class R>
bool traceAny(const orc_scene& s, V3<R> o, V3<R> d, R tmin, R tmax, bool brute) {
  RayW<R> r;
  r.o = o;
  r.d = d;
  r.setup();
  R t, u, v;
  if (brute || s.nodes.empty()) {
    for (uint32_t g = 0; g < s.triInst.size(); ++g)
      if (intersectTri(r, &s.tri[size_t(g) * 9], tmin, tmax, t, u, v)) return true;
    return false;
  }
  V3<R> inv(R(1) / d.x, R(1) / d.y, R(1) / d.z);
  uint32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const orc_scene::Node& n = s.nodes[stack[--sp]];
    if (!hitBox(n, o, inv, tmin, tmax)) continue;
    if (n.count) {
      for (uint32_t i = 0; i < n.count; ++i)
        if (intersectTri(r, &s.tri[size_t(s.order[n.left + i]) * 9], tmin, tmax, t, u, v)) return true;
    } else {
      stack[sp++] = n.left;
      stack[sp++] = n.left + 1;
    }
  }
  return false;
}

// ------------------------------------------------------------ hit state
// HitState + getHitState + configureShadingFrame, rchit_layouts.glsl:37-95.
template <class R> struct HitState {
  int lightId;
  R uvx, uvy;
  V3<R> pos, V, N, ffN, X, Y;
  GpuMaterial mat;
  const float* w2o;  // gl_WorldToObjectEXT of the instance (row-major 3x4)
};

template <class R>
HitState<R> getHitState(const orc_scene& s, const HitR& h, V3<R> rayDir, uint32_t useFaceNormal) {
  HitState<R> st;
  const orc_scene::Inst& in = s.inst[s.triInst[h.tri]];
  uint32_t prim = uint32_t(s.triPrim[h.tri]);
  const std::vector<uint32_t>& ix = s.idx[in.mesh];
  const std::vector<GpuVertex>& vx = s.verts[in.mesh];
  const GpuVertex& v0 = vx[ix[prim * 3 + 0]];
  const GpuVertex& v1 = vx[ix[prim * 3 + 1]];
  const GpuVertex& v2 = vx[ix[prim * 3 + 2]];
  R bx = R(h.u), by = R(h.v);
  V3<R> ba(R(1) - bx - by, bx, by);
  st.lightId = in.light;
  // barymix2 / barymix3, math.glsl:80-86
  st.uvx = R(v0.uv[0]) * ba.x + R(v1.uv[0]) * ba.y + R(v2.uv[0]) * ba.z;
  st.uvy = R(v0.uv[1]) * ba.x + R(v1.uv[1]) * ba.y + R(v2.uv[1]) * ba.z;
  V3<R> op = load3<R>(v0.pos) * ba.x + load3<R>(v1.pos) * ba.y + load3<R>(v2.pos) * ba.z;
  const float* m = in.o2w;
  st.pos = V3<R>(R(m[0]) * op.x + R(m[1]) * op.y + R(m[2]) * op.z + R(m[3]),
                 R(m[4]) * op.x + R(m[5]) * op.y + R(m[6]) * op.z + R(m[7]),
                 R(m[8]) * op.x + R(m[9]) * op.y + R(m[10]) * op.z + R(m[11]));
  V3<R> n = load3<R>(v0.normal) * ba.x + load3<R>(v1.normal) * ba.y + load3<R>(v2.normal) * ba.z;
  const float* w = in.w2o;  // row vector times 3x4: result[c] = sum_r n[r] * w[r][c]
  V3<R> nw(n.x * R(w[0]) + n.y * R(w[4]) + n.z * R(w[8]), n.x * R(w[1]) + n.y * R(w[5]) + n.z * R(w[9]),
           n.x * R(w[2]) + n.y * R(w[6]) + n.z * R(w[10]));
  st.N = makeNormal(nw);
  st.w2o = in.w2o;
  st.V = makeNormal(-rayDir);
  if (st.lightId < 0) st.mat = s.mats[in.material];
  // configureShadingFrame: geoN is never assigned in the reference
  // (rchit_layouts.glsl:62); it is read as the zero vector here.
  if (useFaceNormal == 1) st.N = V3<R>(0);
  basis(st.N, st.X, st.Y);
  st.ffN = dot(st.N, st.V) > R(0) ? st.N : -st.N;
  return st;
}

template <class R> V3<R> textureEval3(const orc_scene& s, int id, R u, R v, R out4[4]) {
  tex2D(s.textures[id], u, v, out4);
  return V3<R>(out4[0], out4[1], out4[2]);
}

template <class R> void applyNormalMap(const orc_scene& s, HitState<R>& st, uint32_t useFaceNormal) {
  R t4[4];
  V3<R> c = textureEval3(s, st.mat.normalTextureId, st.uvx, st.uvy, t4);
  V3<R> n = R(2) * c - V3<R>(1);
  st.N = toWorld(st.X, st.Y, st.N, n);
  st.N = makeNormal(st.N);
  if (useFaceNormal == 1) st.N = V3<R>(0);
  basis(st.N, st.X, st.Y);
  st.ffN = dot(st.N, st.V) > R(0) ? st.N : -st.N;
}

// --------------------------------------------------------- light sampling
template <class R> struct LRec {
  V3<R> d, n;
  R dist = 0, pdf = 0;
  uint32_t flags = 0;
};

template <class R> void envUV(const float* envT, V3<R> L, R& u, R& v, R& theta) {
  // pdfEnvmap / evalEnvmap, sample_light.glsl:84-104:
  // L = transformDirection(transpose(envTransform), L)
  float tr[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) tr[c * 4 + r] = envT[r * 4 + c];
  L = transformDirection(tr, L);
  theta = std::acos(clampr(L.y, R(-1), R(1)));
  u = (K<R>::PI + std::atan2(L.z, L.x)) * K<R>::INV_2PI;
  v = theta * K<R>::INV_PI;
}

template <class R> R pdfEnvmap(const orc_scene& s, const float* envT, R resx, R resy, V3<R> L) {
  R u, v, theta;
  envUV(envT, L, u, v, theta);
  R a[4], b[4];
  tex2D(s.cond, u, v, a);
  tex2D(s.marg, R(0), v, b);
  R pdf = a[1] * b[1];
  R sinTheta = std::sin(theta);
  if (sinTheta == R(0)) return 0;
  return (pdf * resx * resy) / (K<R>::TWO_PI * K<R>::PI * sinTheta);
}

template <class R> V3<R> evalEnvmap(const orc_scene& s, const float* envT, R intensity, V3<R> L) {
  R u, v, theta;
  envUV(envT, L, u, v, theta);
  R e[4];
  tex2D(s.env, u, v, e);
  return intensity * V3<R>(e[0], e[1], e[2]);
}

template <class R>
V3<R> sampleEnvmap(const orc_scene& s, R r1, R r2, const float* envT, R resx, R resy, R intensity,
                   V3<R>& L, R& pdf) {
  // sample_light.glsl:106-129
  R a[4], b[4], c[4], d[4];
  R uvx, uvy;
  tex2D(s.marg, R(0), r1, a);
  uvy = a[0];
  tex2D(s.cond, r2, uvy, b);
  uvx = b[0];
  tex2D(s.cond, uvx, uvy, c);
  tex2D(s.marg, R(0), uvy, d);
  pdf = c[1] * d[1];
  R phi = uvx * K<R>::TWO_PI;
  R theta = uvy * K<R>::PI;
  if (std::sin(theta) == R(0)) pdf = 0;
  pdf = (pdf * resx * resy) / (K<R>::TWO_PI * K<R>::PI * std::sin(theta));
  L = V3<R>(-std::sin(theta) * std::cos(phi), std::cos(theta), -std::sin(theta) * std::sin(phi));
  L = transformDirection(envT, L);
  R e[4];
  tex2D(s.env, uvx, uvy, e);
  return intensity * V3<R>(e[0], e[1], e[2]);
}

template <class R> V3<R> lightU(const GpuLight& l) { return load3<R>(l.u); }
template <class R> V3<R> lightV(const GpuLight& l) { return load3<R>(l.v); }

template <class R> V3<R> sampleOneLight(R r0, R r1, const GpuLight& light, V3<R> p, LRec<R>& l) {
  // sample_light.glsl:10-82
  int type = light.type;
  if (type == LightTypeRect || type == LightTypeTriangle) {
    R a = r0, b = (type == LightTypeRect) ? r1 : (R(1) - r0) * r1;
    V3<R> lp = load3<R>(light.position) + lightU<R>(light) * a + lightV<R>(light) * b;
    l.d = lp - p;
    l.dist = length(l.d);
    R distSq = l.dist * l.dist;
    l.d = l.d / l.dist;
    l.n = makeNormal(cross(lightU<R>(light), lightV<R>(light)));
    l.pdf = distSq / (R(light.area) * std::fabs(dot(l.n, l.d)) + K<R>::EPS);
    l.flags = EArea;
    return load3<R>(light.radiance);
  } else if (type == LightTypeDirectional) {
    l.d = makeNormal(load3<R>(light.direction));
    l.n = -l.d;
    l.dist = K<R>::INF;
    l.pdf = 1;
    l.flags = EDelta;
    return load3<R>(light.radiance);
  } else if (type == LightTypePoint) {
    l.d = load3<R>(light.position) - p;
    l.n = -l.d;
    l.dist = length(l.d);
    R distSq = l.dist * l.dist;
    l.d = l.d / (l.dist + K<R>::EPS);
    V3<R> rad = load3<R>(light.radiance) / (distSq + K<R>::EPS);
    l.pdf = 1;
    l.flags = EDelta;
    return rad;
  }
  // Undefined type: the GLSL function falls off its end; the record keeps
  // whatever it held (zero-initialised here).
  return V3<R>(0);
}

struct RCtx {
  const orc_scene* s;
  const GpuCamera* cam;
  const GpuPushConstantRaytrace* pc;
};

template <class R> struct Payload {
  // PathRecord, BsdfSamplingRecord, MultiChannelRecord, DirectLightRecord
  // (structs.glsl:72-90)
  V3<R> ro, rd, radiance, throughput;
  uint32_t depth = 1, seed = 0;
  bool stop = false;
  V3<R> bd;
  R bpdf = 0;
  uint32_t bflags = 0;
  V3<R> ch[8];
  V3<R> dRad;
  R dDist = 0;
  V3<R> dO, dD;
  bool dSkip = true;
  bool fromDielectric = false;
  R firstT = 0;
};


// ------------------------------------------------------------ sun & sky
// utils/sun_and_sky.glsl:20-533 (NVIDIA's MDL sun & sky: Preetham/Perez
// sky, attenuated sun disk with glow, ground plane below the horizon).  The
// file defines its own M_PI = 3.1415926535f.
template <class R> struct Sky {
  static constexpr R MPI = R(3.1415926535f);
};
template <class R> R skyLum(V3<R> c) { return R(0.2126) * c.x + R(0.7152) * c.y + R(0.0722) * c.z; }
template <class R> R smoothstepR(R e0, R e1, R x) {
  R t = std::fmin(std::fmax((x - e0) / (e1 - e0), R(0)), R(1));
  return t * t * (R(3) - R(2) * t);
}
// xyz2dir, sun_and_sky.glsl:33-57
template <class R> V3<R> skyXyz2dir(V3<R> n, R x, R y, R z) {
  V3<R> u;
  if (std::fabs(n.x) < std::fabs(n.y)) u = V3<R>(0, -n.z, n.y);
  else u = V3<R>(n.z, 0, -n.x);
  u = makeNormal(u);
  V3<R> v = cross(n, u);
  return x * u + y * v + z * n;
}
// mi_lib_square_to_disk + mi_reflection_dir_diffuse_x, sun_and_sky.glsl:59-106
template <class R> V3<R> skyDiffuseDir(V3<R> n, R sx, R sy) {
  R lx = R(2) * sx - R(1), ly = R(2) * sy - R(1), r, phi;
  const R q = Sky<R>::MPI / R(4);
  if (lx == R(0) && ly == R(0)) {
    r = 0, phi = 0;
  } else if (lx > -ly) {
    if (lx > ly) r = lx, phi = q * (R(1) + ly / lx);
    else r = ly, phi = q * (R(3) - lx / ly);
  } else {
    if (lx < ly) r = -lx, phi = q * (R(5) + ly / lx);
    else r = -ly, phi = q * (R(7) - lx / ly);
  }
  R x = r * std::cos(phi), y = r * std::sin(phi);
  R z2 = R(1) - x * x - y * y;
  R z = z2 > R(0) ? std::sqrt(z2) : R(0);
  return skyXyz2dir(n, x, y, z);
}
// calc_sun_color, sun_and_sky.glsl:108-131
template <class R> V3<R> skySunColor(V3<R> sd, R turbidity) {
  V3<R> c(0);
  const R ko[3] = {R(12.0), R(8.5), R(0.9)}, wl[3] = {R(0.610), R(0.550), R(0.470)};
  const R sol[3] = {R(1.0) * R(127500) / R(0.9878), R(0.992) * R(127500) / R(0.9878),
                    R(0.911) * R(127500) / R(0.9878)};
  if (sd.z > R(0)) {
    R m = R(1) / (sd.z + R(0.15) * std::pow(R(93.885) - std::acos(sd.z) * R(180) / Sky<R>::MPI, R(-1.253)));
    R beta = R(0.04608) * turbidity - R(0.04586), alpha = R(1.3), l = R(0.0035);
    R out[3];
    for (int k = 0; k < 3; ++k) {
      R ta = std::exp(-m * beta * std::pow(wl[k], -alpha));
      R to = std::exp(-m * ko[k] * l);
      R tr = std::exp(-m * R(0.008735) * std::pow(wl[k], R(-4.08)));
      out[k] = tr * ta * to * sol[k];
    }
    c = V3<R>(out[0], out[1], out[2]);
  }
  return c;
}
// Perez distribution ratio F(theta, gamma) / F(0, theta_s).
template <class R> R skyPerez(R A, R B, R C, R D, R E, R cosTheta, R gamma, R cosGamma, R thetaSun, R cosThetaSun) {
  return ((R(1) + A * std::exp(B / cosTheta)) * (R(1) + C * std::exp(D * gamma) + E * cosGamma * cosGamma)) /
         ((R(1) + A * std::exp(B / R(1))) *
          (R(1) + C * std::exp(D * thetaSun) + E * cosThetaSun * cosThetaSun));
}
// sky_color_xyz, sun_and_sky.glsl:134-199
template <class R> V3<R> skyColorXyz(V3<R> dir, V3<R> sun, R T, R lum) {
  R cg = dot(sun, dir);
  if (cg > R(1)) cg = R(2) - cg;
  R gamma = std::acos(cg), ct = dir.z, cts = sun.z, ts = std::acos(cts);
  R t2 = T * T, ts2 = ts * ts, ts3 = ts2 * ts;
  R zx = (R(0.001650) * ts3 - R(0.003742) * ts2 + R(0.002088) * ts + R(0)) * t2 +
         (R(-0.029028) * ts3 + R(0.063773) * ts2 - R(0.032020) * ts + R(0.003948)) * T +
         (R(0.116936) * ts3 - R(0.211960) * ts2 + R(0.060523) * ts + R(0.258852));
  R zy = (R(0.002759) * ts3 - R(0.006105) * ts2 + R(0.003162) * ts + R(0)) * t2 +
         (R(-0.042149) * ts3 + R(0.089701) * ts2 - R(0.041536) * ts + R(0.005158)) * T +
         (R(0.153467) * ts3 - R(0.267568) * ts2 + R(0.066698) * ts + R(0.266881));
  R x = skyPerez(R(-0.019257) * T - (R(0.29) - std::pow(cts, R(0.5)) * R(0.09)), R(-0.066513) * T + R(0.000818),
                 R(-0.000417) * T + R(0.212479), R(-0.064097) * T - R(0.898875), R(-0.003251) * T + R(0.045178), ct,
                 gamma, cg, ts, cts);
  R y = skyPerez(R(-0.016698) * T - R(0.260787), R(-0.094958) * T + R(0.009213), R(-0.007928) * T + R(0.210230),
                 R(-0.044050) * T - R(1.653694), R(-0.010922) * T + R(0.052919), ct, gamma, cg, ts, cts);
  const R sat = 1;
  x = zx * ((x * sat) + (R(1) - sat));
  y = zy * ((y * sat) + (R(1) - sat));
  return V3<R>((x / y) * lum, lum, ((R(1) - x - y) / y) * lum);
}
// sky_luminance, sun_and_sky.glsl:201-226
template <class R> R skyLuminance(V3<R> dir, V3<R> sun, R T) {
  R cg = dot(sun, dir);
  if (cg < R(0)) cg = 0;
  if (cg > R(1)) cg = R(2) - cg;
  R gamma = std::acos(cg), ct = dir.z, cts = sun.z, ts = std::acos(cts);
  return skyPerez(R(0.178721) * T - R(1.463037), R(-0.355402) * T + R(0.427494), R(-0.022669) * T + R(5.325056),
                  R(0.120647) * T - R(2.577052), R(-0.066967) * T + R(0.370275), ct, gamma, cg, ts, cts);
}
// calc_env_color, sun_and_sky.glsl:228-244
template <class R> V3<R> skyEnvColor(V3<R> sun, V3<R> dir, R T) {
  R ts = std::acos(sun.z);
  R chi = (R(4.0) / R(9.0) - T / R(120.0)) * (Sky<R>::MPI - R(2) * ts);
  R lum = R(1000.0) * ((R(4.0453) * T - R(4.9710)) * std::tan(chi) - R(0.2155) * T + R(2.4192));
  lum *= skyLuminance(dir, sun, T);
  V3<R> X = skyColorXyz(dir, sun, T, lum);
  V3<R> c(R(3.241) * X.x - R(1.537) * X.y - R(0.499) * X.z, R(-0.969) * X.x + R(1.876) * X.y + R(0.042) * X.z,
          R(0.056) * X.x - R(0.204) * X.y + R(1.057) * X.z);
  return c * Sky<R>::MPI;
}
// calc_irrad, sun_and_sky.glsl:246-262: 5x5 stratified cosine directions;
// the loop counters are GLSL floats.
template <class R> V3<R> skyIrrad(V3<R> sun, R haze) {
  V3<R> acc(0);
  for (float u = 1.f / 10.f; u < 1.f; u += 1.f / 5.f)
    for (float v = 1.f / 10.f; v < 1.f; v += 1.f / 5.f)
      acc = acc + skyEnvColor(sun, skyDiffuseDir(V3<R>(0, 0, 1), R(u), R(v)), haze);
  return acc / R(25.0);
}
// tweak_saturation, sun_and_sky.glsl:264-277
template <class R> R skyTweakSaturation(R sat, R haze) {
  R low = std::pow(sat, R(3));
  if (sat <= R(1)) {
    R h = (haze - R(2)) / R(15);
    h = std::fmin(std::fmax(h, R(0)), R(1));
    h = std::pow(h, R(3));
    return sat * (R(1) - h) + low * h;
  }
  return 1;
}
// arch_vectortweak, sun_and_sky.glsl:279-289
template <class R> V3<R> skyVectorTweak(V3<R> d, int yUp, R hh) {
  V3<R> o = yUp == 1 ? V3<R>(d.x, d.z, d.y) : d;
  if (hh != R(0)) {
    o.z -= hh;
    o = makeNormal(o);
  }
  return o;
}
// arch_colortweak, sun_and_sky.glsl:291-313 (its clamp of boosted
// saturation writes a value nothing reads).
template <class R> V3<R> skyColorTweak(V3<R> tint, R sat, R red) {
  R I = skyLum(tint);
  V3<R> o = sat <= R(0) ? V3<R>(I) : tint * sat + V3<R>(I * (R(1) - sat));
  return o * V3<R>(R(1) + red, R(1), R(1) - red);
}
// calc_physical_scale, sun_and_sky.glsl:315-396: (disk scale, glow scale).
template <class R> void skyPhysicalScale(R diskScale, R glow, R diskInt, R& sdScale, R& sgScale) {
  const R PI = Sky<R>::MPI;
  R diskR = R(0.00465) * diskScale, glowR = diskR * R(10);
  R glowInt = glow * ((R(4) * PI) - (R(24) * PI) / (glowR * glowR) + (R(24) * PI) * std::sin(glowR) / (glowR * glowR * glowR));
  R target = diskInt * PI;
  sgScale = 1;
  R maxGlow = R(0.5) * target;
  if (glowInt > maxGlow) {
    sgScale *= maxGlow / glowInt;
    target -= maxGlow;
  } else {
    target -= glowInt;
  }
  R area = R(2) * PI * (R(1) - std::cos(diskR));
  R targetI = target / area;
  R actual = diskInt * R(100) * (R(1) * area) / area;
  sdScale = targetI == R(0) ? R(0) : targetI / actual;
}
// night_brightness_adjustment, sun_and_sky.glsl:398-405
template <class R> R skyNight(V3<R> sd) {
  const R lmt = R(0.30901699437494742410229341718282);
  if (sd.z <= -lmt) return 0;
  R f = (sd.z + lmt) / lmt;
  f *= f;
  f *= f;
  return f;
}
// sun_and_sky, sun_and_sky.glsl:407-531
template <class R> V3<R> sunAndSky(const GpuSunAndSky& ss, V3<R> inDir) {
  R factor = 1, nightFactor = 1;
  V3<R> rgbScale = load3<R>(ss.rgb_unit_conversion);
  R hh = R(ss.horizon_height) / R(10);
  V3<R> dir = skyVectorTweak(inDir, ss.y_is_up, hh);
  R haze = R(2) + R(ss.haze);
  if (haze < R(2)) haze = 2;
  R sat = skyTweakSaturation(R(ss.saturation), haze);
  if (skyLum(rgbScale) < R(0)) rgbScale = V3<R>(R(1) / R(80000));
  rgbScale = rgbScale * R(ss.multiplier);
  if (R(ss.multiplier) <= R(0)) return V3<R>(0);
  R down = dir.z;
  V3<R> realDir = dir;
  if (dir.z < R(0.001)) {
    dir.z = R(0.001);
    dir = makeNormal(dir);
  }
  V3<R> sun = skyVectorTweak(makeNormal(load3<R>(ss.sun_direction)), ss.y_is_up, hh);
  V3<R> realSun = sun;
  if (sun.z < R(0.001)) {
    if (sun.z < R(0)) factor = skyNight(sun);
    sun.z = R(0.001);
    sun = makeNormal(sun);
  }
  V3<R> tint(0);
  if (factor > R(0)) {
    tint = skyEnvColor(sun, dir, haze);
    if (factor < R(1)) tint = tint * factor;
  }
  V3<R> sunColor = skySunColor(sun, down > R(0) ? haze : R(2));
  if (R(ss.sun_disk_intensity) > R(0) && R(ss.sun_disk_scale) > R(0)) {
    R angle = std::acos(dot(realDir, realSun));
    R radius = R(0.00465) * R(ss.sun_disk_scale) * R(10);
    if (angle < radius) {
      R sd = 1, sg = 1;
      if (ss.physically_scaled_sun == 1)
        skyPhysicalScale(R(ss.sun_disk_scale), R(ss.sun_glow_intensity), R(ss.sun_disk_intensity), sd, sg);
      R f = (R(1) - angle / radius) * R(10);
      f = std::pow(f / R(10), R(3)) * R(2) * R(ss.sun_glow_intensity) * sg +
          smoothstepR(R(8.5), R(9.5) + (haze / R(50)), f) * R(100) * R(ss.sun_disk_intensity) * sd;
      tint = tint + sunColor * f;
    }
  }
  V3<R> out = tint * rgbScale;
  if (down <= R(0)) {
    V3<R> irrad = skyIrrad(sun, R(2));
    V3<R> dc = load3<R>(ss.ground_color) * (irrad + sunColor * sun.z) * rgbScale;
    if (factor < R(1)) dc = dc * factor;
    R blur = R(ss.horizon_blur) / R(10);
    if (blur > R(0)) {
      R dn = -down / blur;
      if (dn > R(1)) dn = 1;
      dn = smoothstepR(R(0), R(1), dn);
      out = out * (R(1) - dn) + dc * dn;
      nightFactor = R(1) - dn;
    } else {
      out = dc;
      nightFactor = 0;
    }
  }
  out = skyColorTweak(out, sat, R(ss.redblueshift));
  V3<R> res = out;
  if (nightFactor > R(0)) {
    V3<R> night = load3<R>(ss.night_color) * nightFactor;
    res = V3<R>(std::fmax(res.x, night.x), std::fmax(res.y, night.y), std::fmax(res.z, night.z));
  }
  return res * Sky<R>::MPI;
}

template <class R>
V3<R> sampleLights(const RCtx& c, Payload<R>& pl, V3<R> pos, V3<R> nrm, bool& visible, LRec<R>& l) {
  // rchit_layouts.glsl:118-167
  const GpuPushConstantRaytrace& pc = *c.pc;
  bool allowDoubleSide = false;
  V3<R> radiance(0);
  const bool sky = c.s->sky.in_use == 1;
  bool hasEnv = pc.hasEnvMap == 1 || sky;
  bool hasLight = pc.numLights > 0;
  R envSel, anaSel;
  if (hasEnv && hasLight) envSel = anaSel = R(0.5);
  else if (hasEnv) envSel = 1, anaSel = 0;
  else if (hasLight) envSel = 0, anaSel = 1;
  else envSel = anaSel = 0;
  R sel = rnd<R>(pl.seed);
  if (sel < envSel) {
    // sampleEnvironmentLight, rchit_layouts.glsl:98-116
    l.flags = EArea;
    l.dist = K<R>::INF;
    l.n = -makeNormal(pl.rd);
    R r0 = rnd<R>(pl.seed), r1 = rnd<R>(pl.seed);
    V3<R> L;
    R pdf;
    V3<R> rad;
    if (sky) {  // sampleSunSky, sample_light.glsl:137-141
      L = uniformSampleSphere(r0, r1);
      pdf = K<R>::INV_4PI;
      rad = sunAndSky(c.s->sky, L);
    } else {
      rad = sampleEnvmap(*c.s, r0, r1, c.cam->envTransform, R(pc.envMapResolution[0]), R(pc.envMapResolution[1]),
                         R(pc.envMapIntensity), L, pdf);
    }
    l.d = L;
    l.pdf = pdf;
    radiance = rad / envSel;
    allowDoubleSide = true;
  } else if (sel < envSel + anaSel) {
    int li = std::min(1 + int(rnd<R>(pl.seed) * R(pc.numLights)), pc.numLights);
    const GpuLight& light = c.s->lights[li];
    R r0 = rnd<R>(pl.seed), r1 = rnd<R>(pl.seed);
    radiance = sampleOneLight(r0, r1, light, pos, l) * R(pc.numLights) / anaSel;
    allowDoubleSide = (light.doubleSide == 1);
  }
  pl.dO = offsetPositionAlongNormal(pos, nrm);
  pl.dD = l.d;
  pl.dDist = l.dist;
  visible = (dot(l.d, nrm) > R(0) && l.pdf > R(0));
  visible = visible && (dot(l.n, l.d) < R(0) || allowDoubleSide);
  return radiance;
}

// -------------------------------------------------------------- BxDFs
template <class R> V3<R> mixv(V3<R> x, V3<R> y, R a) { return x * (R(1) - a) + y * a; }
template <class R> R dAnisoGGX(R HdotN, R HdotX, R HdotY, R ax, R ay) {
  return R(1) / (K<R>::PI * ax * ay * sqr(sqr(HdotX / ax) + sqr(HdotY / ay) + sqr(HdotN)) + K<R>::EPS);
}
template <class R> V3<R> importanceSampleAnisoGGX(R u1, R u2, V3<R> wo, R ax, R ay) {
  R factor = safeSqrt(u1 / std::max(R(1) - u1, K<R>::EPS));
  R phi = K<R>::TWO_PI * u2;
  V3<R> wh(0, 0, 1);
  wh.x = -ax * factor * std::cos(phi);
  wh.y = -ay * factor * std::sin(phi);
  wh = makeNormal(wh);
  return reflect(-wo, wh);
}
template <class R> R importanceAnisoGGXPdf(V3<R> wh, V3<R> wo, R ax, R ay) {
  R pdf = 0, HdotV = dot(wh, wo);
  V3<R> wi = reflect(-wo, wh);
  if (wi.z > R(0) && wo.z > R(0) && wh.z > R(0))
    pdf = dAnisoGGX(wh.z, wh.x, wh.y, ax, ay) * std::fabs(wh.z) / (R(4) * HdotV + K<R>::EPS);
  return pdf;
}
template <class R> R g1SmithAnisoGGX(R NdotV, R VdotX, R VdotY, R ax, R ay) {
  if (NdotV <= R(0)) return 0;
  V3<R> f(ax * VdotX, ay * VdotY, NdotV);
  return R(1) / (NdotV + length(f));
}

// --- Lambertian, brdf_lambertian.rchit:12-42
template <class R> V3<R> lamEval(V3<R> L, V3<R> V, V3<R> N, V3<R> diffuse, uint32_t flag) {
  if ((flag & EArea) == 0) return V3<R>(0);
  R NdotL = dot(N, L), NdotV = dot(N, V);
  if (NdotL < R(0) || NdotV < R(0)) return V3<R>(0);
  return diffuse * K<R>::INV_PI * NdotL;
}
template <class R> R lamPdf(V3<R> L, V3<R> N, uint32_t flag) {
  if ((flag & EArea) == 0) return 0;
  return cosineHemispherePdf(dot(N, L));
}

// --- PBR metalness-roughness, brdf_pbr_metalness_roughness.rchit:14-129
template <class R> V3<R> fresnelSchlick(R c, V3<R> f0) {
  R p = std::pow(clampr(R(1) - c, R(0), R(1)), R(5));
  return f0 + (V3<R>(1) - f0) * p;
}
template <class R>
V3<R> pbrEval(V3<R> L, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, V3<R> albedo, R ax, R ay, R metal, R eta) {
  V3<R> weight(0);
  R NdotL = dot(N, L), NdotV = dot(N, V);
  if (NdotL <= R(0) || NdotV <= R(0)) return weight;
  V3<R> H = makeNormal(L + V);
  R HdotN = dot(H, N), HdotX = dot(H, X), HdotY = dot(H, Y), HdotV = dot(H, V);
  R F0 = sqr((eta - R(1)) / (eta + R(1)));
  V3<R> Fs = fresnelSchlick(HdotV, mixv(V3<R>(F0), albedo, metal));
  R Ds = dAnisoGGX(HdotN, HdotX, HdotY, ax, ay);
  R Gs = g1SmithAnisoGGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
  Gs *= g1SmithAnisoGGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
  V3<R> diffuse = albedo * (R(1) - metal) * K<R>::INV_PI;
  V3<R> specular = Fs * Ds * Gs;
  weight = weight + diffuse;
  weight = weight + specular;
  weight = weight * NdotL;
  return weight;
}
template <class R> R pbrPdf(V3<R> L, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, R ax, R ay, uint32_t flags) {
  R pdf = 0;
  if ((flags & EArea) == 0) return pdf;
  R NdotL = dot(N, L), NdotV = dot(N, V);
  if (NdotL <= R(0) || NdotV <= R(0)) return pdf;
  V3<R> H = makeNormal(L + V);
  V3<R> wh = makeNormal(toLocal(X, Y, N, H));
  V3<R> wo = makeNormal(toLocal(X, Y, N, V));
  V3<R> wi = makeNormal(toLocal(X, Y, N, L));
  pdf += R(0.2) * cosineHemispherePdf(wi.z);
  pdf += (R(1) - R(0.2)) * importanceAnisoGGXPdf(wh, wo, ax, ay);
  return pdf;
}
template <class R>
V3<R> pbrSample(R u0, R u1, uint32_t& seed, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, V3<R> albedo, R ax,
                R ay, R metal, R eta, V3<R>& d, R& pdf, uint32_t& flags) {
  V3<R> wo = makeNormal(toLocal(X, Y, N, V));
  V3<R> wi;
  pdf = 0;
  if (rnd<R>(seed) < R(0.2)) {
    wi = cosineSampleHemisphere(u0, u1);
    flags = EDiffuseReflection;
    pdf = cosineHemispherePdf(wi.z);
  } else {
    wi = importanceSampleAnisoGGX(u0, u1, wo, ax, ay);
    V3<R> wh = makeNormal(wi + wo);
    pdf = importanceAnisoGGXPdf(wh, wo, ax, ay);
    flags = EGlossyReflection;
  }
  d = toWorld(X, Y, N, wi);
  return pbrEval(d, V, N, X, Y, albedo, ax, ay, metal, eta) * std::fabs(wi.z);
}

// --- Rough plastic, brdf_rough_plastic.rchit:12-189
template <class R> R fresnelDielectricExt(R cosThetaI_, R eta) {
  if (eta == R(1)) return 0;
  R scale = (cosThetaI_ > R(0)) ? R(1) / eta : eta;
  R cosThetaTSqr = R(1) - (R(1) - cosThetaI_ * cosThetaI_) * (scale * scale);
  if (cosThetaTSqr <= R(0)) return 1;
  R cosThetaI = std::fabs(cosThetaI_);
  R cosThetaT = std::sqrt(cosThetaTSqr);
  R Rs = (cosThetaI - eta * cosThetaT) / (cosThetaI + eta * cosThetaT);
  R Rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
  return R(0.5) * (Rs * Rs + Rp * Rp);
}
// eval(L,V,N,X,Y,kd,ks,eta,fdrInt,invEta2,ax,ay,flags) -- signature order.
template <class R>
V3<R> rpEval(V3<R> L, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, V3<R> kd, V3<R> ks, R eta, R fdrInt,
             R invEta2, R ax, R ay, uint32_t flags) {
  V3<R> weight(0);
  R NdotL = dot(L, N), NdotV = dot(V, N);
  if (((flags & EArea) == 0) || NdotL < R(0) || NdotV < R(0)) return weight;
  {
    V3<R> H = makeNormal(L + V);
    R Fs = fresnelDielectricExt(dot(H, V), eta);
    R Gs = g1SmithAnisoGGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
    Gs *= g1SmithAnisoGGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
    R Ds = dAnisoGGX(dot(H, N), dot(H, X), dot(H, Y), ax, ay);
    weight = weight + ks * Fs * Gs * Ds * NdotL;
  }
  {
    R Fo = fresnelDielectricExt(NdotV, eta);
    R Fi = fresnelDielectricExt(NdotL, eta);
    V3<R> diff = kd;
    diff = diff / (V3<R>(1) - diff * fdrInt);
    weight = weight + (R(1) - Fi) * (R(1) - Fo) * diff * invEta2 * K<R>::INV_PI * NdotL;
  }
  return weight;
}
// pdf(L,V,N,X,Y,ax,ay,eta,ssw,flags) -- signature order.
template <class R>
R rpPdf(V3<R> L, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, R ax, R ay, R eta, R ssw, uint32_t flags) {
  R pdf = 0;
  R NdotL = dot(L, N), NdotV = dot(V, N);
  if (NdotL < R(0) || NdotV < R(0) || ((flags & EArea) == 0)) return pdf;
  R Fo = fresnelDielectricExt(NdotV, eta);
  R substrateWeight = ssw * (R(1) - Fo);
  R specularWeight = Fo;
  R probSpecular = specularWeight / (specularWeight + substrateWeight);
  V3<R> wi = toLocal(X, Y, N, L), wo = toLocal(X, Y, N, V);
  V3<R> wh = makeNormal(wi + wo);
  R glossyPdf = importanceAnisoGGXPdf(wh, wo, ax, ay);
  pdf += probSpecular * glossyPdf;
  R diffusePdf = cosineHemispherePdf(wi.z);
  pdf += (R(1) - probSpecular) * diffusePdf;
  return pdf;
}
template <class R>
V3<R> rpSample(R u0, R u1, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, V3<R> kd, V3<R> ks, R eta, R fdrInt,
               R invEta2, R ax, R ay, R ssw, V3<R>& d, R& pdf, uint32_t& flags) {
  V3<R> weight(0);
  R NdotV = dot(V, N);
  if (NdotV <= R(0)) {
    flags = EBsdfNull;
    pdf = 0;
    d = V3<R>(0);
    return weight;
  }
  R Fo = fresnelDielectricExt(NdotV, eta);
  R substrateWeight = ssw * (R(1) - Fo);
  R specularWeight = Fo;
  R probSpecular = specularWeight / (specularWeight + substrateWeight);
  if (u0 < probSpecular) {
    u0 = u0 / probSpecular;
    V3<R> wo = toLocal(X, Y, N, V);
    V3<R> wi = importanceSampleAnisoGGX(u0, u1, wo, ax, ay);
    V3<R> L = d = toWorld(X, Y, N, wi);
    V3<R> H = makeNormal(V + L);
    V3<R> wh = makeNormal(wi + wo);
    R NdotL = dot(N, L);
    R Fs = fresnelDielectricExt(dot(H, V), eta);
    R Gs = g1SmithAnisoGGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
    Gs *= g1SmithAnisoGGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
    R Ds = dAnisoGGX(dot(H, N), dot(H, X), dot(H, Y), ax, ay);
    flags = EGlossyReflection;
    pdf = importanceAnisoGGXPdf(wh, wo, ax, ay) * probSpecular;
    weight = ks * Fs * Gs * Ds * NdotL;
  } else {
    u0 = (u0 - probSpecular) / (R(1) + K<R>::EPS - probSpecular);
    V3<R> wi = cosineSampleHemisphere(u0, u1);
    R Fi = fresnelDielectricExt(wi.z, eta);
    V3<R> diff = kd;
    diff = diff / (V3<R>(1) - diff * fdrInt);
    pdf = (R(1) - probSpecular) * cosineHemispherePdf(wi.z);
    d = toWorld(X, Y, N, wi);
    flags = EDiffuseReflection;
    weight = (R(1) - Fi) * (R(1) - Fo) * invEta2 * diff * K<R>::INV_PI * std::fabs(wi.z);
  }
  return weight;
}

// --- Dielectric, bsdf_dielectric.rchit:12-138
template <class R> R dielectricFresnel(R cosThetaI, R eta) {
  R sinThetaTSq = eta * eta * (R(1) - cosThetaI * cosThetaI);
  if (sinThetaTSq > R(1)) return 1;
  R cosThetaT = std::sqrt(std::max(R(1) - sinThetaTSq, R(0)));
  R rs = (eta * cosThetaT - cosThetaI) / (eta * cosThetaT + cosThetaI);
  R rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
  return R(0.5) * (rs * rs + rp * rp);
}
template <class R> V3<R> dielEval(V3<R> L, V3<R> V, V3<R> N, R eta, uint32_t flags) {
  V3<R> weight(0);
  if ((flags & EDelta) == 0) return weight;
  R F = dielectricFresnel(std::fabs(dot(V, N)), eta);
  if (dot(L, N) > R(0)) {
    if (std::fabs(dot(reflect(-V, N), L) - R(1)) < K<R>::EPS) weight = weight + F * V3<R>(1);
  } else {
    if (std::fabs(dot(refract(-V, N, eta), L) - R(1)) < K<R>::EPS) weight = weight + (R(1) - F) * V3<R>(1) * eta * eta;
  }
  return weight;
}
template <class R> R dielPdf(V3<R> L, V3<R> V, V3<R> N, R eta, uint32_t flags) {
  R pdf = 0;
  if ((flags & EDelta) == 0) return pdf;
  R F = dielectricFresnel(std::fabs(dot(V, N)), eta);
  if (dot(L, N) > R(0)) {
    if (std::fabs(dot(reflect(-V, N), L) - R(1)) < K<R>::EPS) pdf = F;
  } else {
    if (std::fabs(dot(refract(-V, N, eta), L) - R(1)) < K<R>::EPS) pdf = R(1) - F;
  }
  return pdf;
}
template <class R>
V3<R> dielSample(R u, V3<R> V, V3<R> N, R eta, V3<R>& d, R& pdf, uint32_t& flags) {
  R F = dielectricFresnel(std::fabs(dot(V, N)), eta);
  if (u < F) {
    d = makeNormal(reflect(-V, N));
    pdf = F;
    flags = ESpecularReflection;
    return F * V3<R>(1);
  }
  d = makeNormal(refract(-V, N, eta));
  pdf = R(1) - F;
  flags = ESpecularTransmission;
  return (R(1) - F) * V3<R>(1) * eta * eta;
}

template <class R> R luminance(V3<R> c) {  // sun_and_sky.glsl:29-31
  return R(0.2126f) * c.x + R(0.7152f) * c.y + R(0.0722f) * c.z;
}

// ------------------------------------------------------ closest-hit shaders
template <class R> void writeAovLambertian(const RCtx& c, Payload<R>& pl, const HitState<R>& st) {
  // brdf_lambertian.rchit:93-108
  const GpuPushConstantRaytrace& pc = *c.pc;
  if (pc.diffuseOutChannel >= 0) pl.ch[pc.diffuseOutChannel] = load3<R>(st.mat.diffuse);
  if (pc.normalOutChannel >= 0) pl.ch[pc.normalOutChannel] = st.N;
  if (pc.specularOutChannel >= 0) pl.ch[pc.specularOutChannel] = V3<R>(0);
  if (pc.tangentOutChannel >= 0) pl.ch[pc.tangentOutChannel] = st.X;
  if (pc.roughnessOutChannel >= 0) pl.ch[pc.roughnessOutChannel] = V3<R>(1, 1, 0);
  if (pc.positionOutChannel >= 0) pl.ch[pc.positionOutChannel] = st.pos;
  if (pc.uvOutChannel >= 0) pl.ch[pc.uvOutChannel] = V3<R>(st.uvx, st.uvy, 1);
}

template <class R> void chsLambertian(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  const orc_scene& s = *c.s;
  const GpuPushConstantRaytrace& pc = *c.pc;
  if (st.lightId >= 0) {
    // hitLight, brdf_lambertian.rchit:44-68
    const GpuLight& light = s.lights[st.lightId];
    V3<R> lightDirection = makeNormal(pl.rd);
    V3<R> lightNormal = makeNormal(cross(lightU<R>(light), lightV<R>(light)));
    R proj = dot(lightNormal, lightDirection);
    pl.stop = true;
    if (proj > R(0) && light.doubleSide == 0) return;
    R mis = 1;
    if (isNonSpecular(pl.bflags) && pl.depth != 1) {
      R lightDist = length(st.pos - pl.ro);
      R distSquare = lightDist * lightDist;
      R lightPdf = distSquare / (R(light.area) * std::fabs(proj) + K<R>::EPS);
      mis = powerHeuristic(pl.bpdf, lightPdf);
    }
    pl.radiance = pl.radiance + pl.throughput * load3<R>(light.radiance) * mis;
    return;
  }
  R t4[4];
  if (st.mat.diffuseTextureId >= 0) {
    V3<R> d = textureEval3(s, st.mat.diffuseTextureId, st.uvx, st.uvy, t4);
    st.mat.diffuse[0] = float(d.x), st.mat.diffuse[1] = float(d.y), st.mat.diffuse[2] = float(d.z);
  }
  if (st.mat.normalTextureId >= 0) applyNormalMap(s, st, pc.useFaceNormal);
  V3<R> kd = load3<R>(st.mat.diffuse);
  if (pl.depth == 1) writeAovLambertian(c, pl, st);
  {
    V3<R> Ld(0);
    bool visible;
    LRec<R> l;
    V3<R> radiance = sampleLights(c, pl, st.pos, st.ffN, visible, l);
    if (visible) {
      V3<R> bw = lamEval(l.d, st.V, st.ffN, kd, EArea);
      R bsdfPdf = lamPdf(l.d, st.ffN, l.flags);
      R mis = powerHeuristic(l.pdf, bsdfPdf);
      Ld = Ld + mis * bw * radiance * pl.throughput / (l.pdf + K<R>::EPS);
    }
    pl.dRad = Ld;
    pl.dSkip = !visible;
  }
  R u0 = rnd<R>(pl.seed), u1 = rnd<R>(pl.seed);
  V3<R> wi = cosineSampleHemisphere(u0, u1);
  R bpdf = cosineHemispherePdf(wi.z);
  V3<R> bd = toWorld(st.X, st.Y, st.ffN, wi);
  V3<R> bw = kd * K<R>::INV_PI * std::fabs(wi.z);
  if (bpdf <= R(0) || length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  pl.bd = bd;
  pl.bpdf = bpdf;
  pl.bflags = EDiffuseReflection;
  pl.ro = offsetPositionAlongNormal(st.pos, st.ffN);
  pl.rd = bd;
  pl.throughput = pl.throughput * (bw / bpdf);
}

template <class R> void chsPbr(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  const orc_scene& s = *c.s;
  const GpuPushConstantRaytrace& pc = *c.pc;
  R t4[4];
  if (st.mat.diffuseTextureId >= 0) {
    V3<R> d = textureEval3(s, st.mat.diffuseTextureId, st.uvx, st.uvy, t4);
    st.mat.diffuse[0] = float(d.x), st.mat.diffuse[1] = float(d.y), st.mat.diffuse[2] = float(d.z);
  }
  R metal = R(st.mat.metalness), rough = R(st.mat.roughness);
  if (st.mat.metalnessTextureId >= 0) {
    textureEval3(s, st.mat.metalnessTextureId, st.uvx, st.uvy, t4);
    metal = t4[0];
  }
  if (st.mat.roughnessTextureId >= 0) {
    textureEval3(s, st.mat.roughnessTextureId, st.uvx, st.uvy, t4);
    rough = t4[0];
  }
  if (st.mat.normalTextureId >= 0) applyNormalMap(s, st, pc.useFaceNormal);
  R opacity = R(st.mat.specular);
  if (st.mat.opacityTextureId >= 0) {
    textureEval3(s, st.mat.opacityTextureId, st.uvx, st.uvy, t4);
    opacity = t4[0];
  }
  if (rnd<R>(pl.seed) < opacity) {
    pl.ro = offsetPositionAlongNormal(st.pos, -st.ffN);
    pl.depth--;
    return;
  }
  R ax = std::max(sqr(rough), R(0.001));
  R ay = ax;
  R eta = R(st.mat.ior);
  V3<R> kd = load3<R>(st.mat.diffuse);
  if (pl.depth == 1) {
    if (pc.diffuseOutChannel >= 0) pl.ch[pc.diffuseOutChannel] = kd;
    if (pc.normalOutChannel >= 0) pl.ch[pc.normalOutChannel] = st.ffN;
  }
  {
    V3<R> Ld(0);
    bool visible;
    LRec<R> l;
    V3<R> radiance = sampleLights(c, pl, st.pos, st.ffN, visible, l);
    if (visible) {
      V3<R> bw = pbrEval(l.d, st.V, st.ffN, st.X, st.Y, kd, ax, ay, metal, eta);
      R bsdfPdf = pbrPdf(l.d, st.V, st.ffN, st.X, st.Y, ax, ay, l.flags);
      R mis = powerHeuristic(l.pdf, bsdfPdf);
      Ld = mis * bw * radiance * pl.throughput / (l.pdf + K<R>::EPS);
    }
    pl.dRad = Ld;
    pl.dSkip = !visible;
  }
  R u0 = rnd<R>(pl.seed), u1 = rnd<R>(pl.seed);
  V3<R> bd;
  R bpdf;
  uint32_t bflags;
  V3<R> bw = pbrSample(u0, u1, pl.seed, st.V, st.ffN, st.X, st.Y, kd, ax, ay, metal, eta, bd, bpdf, bflags);
  if (bpdf <= R(0) || length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  pl.bd = bd;
  pl.bpdf = bpdf;
  pl.bflags = bflags;
  pl.ro = offsetPositionAlongNormal(st.pos, st.ffN);
  pl.rd = bd;
  pl.throughput = pl.throughput * (bw / (bpdf + K<R>::EPS));
}

// The light-sample terms of brdf_rough_plastic.rchit:229-241: eval and pdf
// are called with the shader's permuted arguments (eval binds
// invEta2 <- ax, ax <- ay, ay <- invEta2; pdf binds ax <- eta, ay <- ax,
// eta <- ay, against the signatures at rchit:82-83 and 113-114).
template <class R>
void rpLightTerms(V3<R> L, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, V3<R> kd, R eta, R fdrInt, R invEta2, R ax, R ay,
                  R ssw, uint32_t lightFlags, V3<R>& bw, R& bsdfPdf) {
  bw = rpEval(L, V, N, X, Y, kd, V3<R>(1), eta, fdrInt, /*invEta2<-*/ ax, /*ax<-*/ ay, /*ay<-*/ invEta2, EArea);
  bsdfPdf = rpPdf(L, V, N, X, Y, /*ax<-*/ eta, /*ay<-*/ ax, /*eta<-*/ ay, ssw, lightFlags);
}

template <class R> void chsRoughPlastic(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  const orc_scene& s = *c.s;
  const GpuPushConstantRaytrace& pc = *c.pc;
  R t4[4];
  if (st.mat.diffuseTextureId >= 0) {
    V3<R> d = textureEval3(s, st.mat.diffuseTextureId, st.uvx, st.uvy, t4);
    st.mat.diffuse[0] = float(d.x), st.mat.diffuse[1] = float(d.y), st.mat.diffuse[2] = float(d.z);
  }
  R alx = R(st.mat.anisoAlpha[0]), aly = R(st.mat.anisoAlpha[1]);
  if (st.mat.roughnessTextureId >= 0) {
    textureEval3(s, st.mat.roughnessTextureId, st.uvx, st.uvy, t4);
    alx = t4[0];
    aly = t4[1];
  }
  if (st.mat.normalTextureId >= 0) applyNormalMap(s, st, pc.useFaceNormal);
  R eta = R(st.mat.ior);
  R fdrInt = R(st.mat.radiance[0]);
  R invEta2 = R(1) / (eta * eta);
  R ax = std::max(K<R>::EPS, alx), ay = std::max(K<R>::EPS, aly);
  V3<R> kd = load3<R>(st.mat.diffuse);
  R ssw = luminance(kd);
  V3<R> ks(1);
  {
    V3<R> Ld(0);
    bool visible;
    LRec<R> l;
    V3<R> radiance = sampleLights(c, pl, st.pos, st.ffN, visible, l);
    if (visible) {
      V3<R> bw;
      R bsdfPdf;
      rpLightTerms(l.d, st.V, st.ffN, st.X, st.Y, kd, eta, fdrInt, invEta2, ax, ay, ssw, l.flags, bw, bsdfPdf);
      R mis = powerHeuristic(l.pdf, bsdfPdf);
      Ld = Ld + mis * bw * radiance * pl.throughput / (l.pdf + K<R>::EPS);
    }
    pl.dRad = Ld;
    pl.dSkip = !visible;
  }
  R u0 = rnd<R>(pl.seed), u1 = rnd<R>(pl.seed);
  V3<R> bd;
  R bpdf;
  uint32_t bflags;
  V3<R> bw = rpSample(u0, u1, st.V, st.ffN, st.X, st.Y, kd, ks, eta, fdrInt, invEta2, ax, ay, ssw, bd, bpdf, bflags);
  if (bpdf <= R(0) || length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  pl.bd = bd;
  pl.bpdf = bpdf;
  pl.bflags = bflags;
  pl.ro = offsetPositionAlongNormal(st.pos, st.ffN);
  pl.rd = bd;
  pl.throughput = pl.throughput * (bw / (bpdf + K<R>::EPS));
}

template <class R> void chsDielectric(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  const orc_scene& s = *c.s;
  const GpuPushConstantRaytrace& pc = *c.pc;
  if (st.mat.normalTextureId >= 0) applyNormalMap(s, st, pc.useFaceNormal);
  R ior = R(st.mat.ior);
  R eta = dot(st.V, st.N) > R(0) ? (R(1) / ior) : ior;
  {
    V3<R> Ld(0);
    bool visible;
    LRec<R> l;
    V3<R> radiance = sampleLights(c, pl, st.pos, st.ffN, visible, l);
    if (visible) {
      V3<R> bw = dielEval(l.d, st.V, st.ffN, eta, EArea);
      R bsdfPdf = dielPdf(l.d, st.V, st.ffN, eta, l.flags);
      R mis = powerHeuristic(l.pdf, bsdfPdf);
      Ld = Ld + mis * bw * radiance * pl.throughput / (l.pdf + K<R>::EPS);
    }
    pl.dRad = Ld;
    pl.dSkip = !visible;
    pl.fromDielectric = true;
  }
  R u = rnd<R>(pl.seed);
  V3<R> bd;
  R bpdf;
  uint32_t bflags;
  V3<R> bw = dielSample(u, st.V, st.ffN, eta, bd, bpdf, bflags);
  if (bpdf <= R(0) || length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  pl.bd = bd;
  pl.bpdf = bpdf;
  pl.bflags = bflags;
  R sg = dot(bd, st.N);
  R sgn = sg > R(0) ? R(1) : (sg < R(0) ? R(-1) : R(0));
  pl.ro = offsetPositionAlongNormal(st.pos, sgn * st.N);
  pl.rd = bd;
  pl.throughput = pl.throughput * (bw / (bpdf + K<R>::EPS));
}

// ---------------------------------------------- further plugins of the set
// brdf_conductor.rchit:14-34 (Shirley's conductor Fresnel; identical in
// brdf_rough_conductor.rchit:14-34)
template <class R> R conductorReflectance(R eta, R k, R cosThetaI) {
  R cosThetaISq = cosThetaI * cosThetaI;
  R sinThetaISq = std::max(R(1) - cosThetaISq, R(0));
  R sinThetaIQu = sinThetaISq * sinThetaISq;
  R innerTerm = eta * eta - k * k - sinThetaISq;
  R aSqPlusBSq = std::sqrt(std::max(innerTerm * innerTerm + R(4) * eta * eta * k * k, R(0)));
  R a = std::sqrt(std::max((aSqPlusBSq + innerTerm) * R(0.5), R(0)));
  R Rs = ((aSqPlusBSq + cosThetaISq) - (R(2) * a * cosThetaI)) / ((aSqPlusBSq + cosThetaISq) + (R(2) * a * cosThetaI));
  R Rp = ((cosThetaISq * aSqPlusBSq + sinThetaIQu) - (R(2) * a * cosThetaI * sinThetaISq)) /
         ((cosThetaISq * aSqPlusBSq + sinThetaIQu) + (R(2) * a * cosThetaI * sinThetaISq));
  return R(0.5) * (Rs + Rs * Rp);
}
template <class R> V3<R> conductorReflectance3(V3<R> eta, V3<R> k, R c) {
  return V3<R>(conductorReflectance(eta.x, k.x, c), conductorReflectance(eta.y, k.y, c),
               conductorReflectance(eta.z, k.z, c));
}

// Next-event estimate with a BSDF value / pdf pair (the USE_MIS block every
// closest-hit shader repeats).
template <class R, class EvalF, class PdfF>
void directLight(const RCtx& c, Payload<R>& pl, HitState<R>& st, EvalF eval, PdfF pdf) {
  V3<R> Ld(0);
  bool visible;
  LRec<R> l;
  V3<R> radiance = sampleLights(c, pl, st.pos, st.ffN, visible, l);
  if (visible) {
    V3<R> bw = eval(l);
    R bsdfPdf = pdf(l);
    R mis = powerHeuristic(l.pdf, bsdfPdf);
    Ld = Ld + mis * bw * radiance * pl.throughput / (l.pdf + K<R>::EPS);
  }
  pl.dRad = Ld;
  pl.dSkip = !visible;
}

template <class R> void diffuseAndNormalMaps(const RCtx& c, HitState<R>& st) {
  R t4[4];
  if (st.mat.diffuseTextureId >= 0) {
    V3<R> d = textureEval3(*c.s, st.mat.diffuseTextureId, st.uvx, st.uvy, t4);
    st.mat.diffuse[0] = float(d.x), st.mat.diffuse[1] = float(d.y), st.mat.diffuse[2] = float(d.z);
  }
}

// Continue the path along a sampled direction (offset along ffN).
template <class R>
void nextRay(Payload<R>& pl, HitState<R>& st, V3<R> bd, V3<R> bw, R bpdf, uint32_t bflags, bool epsInDivisor) {
  pl.bd = bd;
  pl.bpdf = bpdf;
  pl.bflags = bflags;
  pl.ro = offsetPositionAlongNormal(st.pos, st.ffN);
  pl.rd = bd;
  pl.throughput = pl.throughput * (epsInDivisor ? bw / (bpdf + K<R>::EPS) : bw / bpdf);
}

// brdf_mirror.rchit:12-101.  Its NEE term is always zero (eval tests EDelta
// and is called with EArea); the shadow ray is still issued when visible.
template <class R> void chsMirror(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  diffuseAndNormalMaps(c, st);
  if (st.mat.normalTextureId >= 0) applyNormalMap(*c.s, st, c.pc->useFaceNormal);
  V3<R> kd = load3<R>(st.mat.diffuse);
  directLight(
      c, pl, st, [&](const LRec<R>&) { return V3<R>(0); },  // eval(..., EArea)
      [&](const LRec<R>& l) {
        if ((l.flags & EDelta) == 0) return R(0);
        return std::fabs(dot(reflect(-st.V, st.ffN), l.d) - R(1)) < K<R>::EPS ? R(1) : R(0);
      });
  pl.fromDielectric = true;
  V3<R> bd = reflect(-st.V, st.ffN);
  if (length(kd) == R(0)) {
    pl.stop = true;
    return;
  }
  nextRay(pl, st, bd, kd, R(1), ESpecularReflection, false);
}

// brdf_conductor.rchit:36-155.  NEE is always zero as for the mirror; the
// sample consumes two draws it does not use.
template <class R> void chsConductor(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  diffuseAndNormalMaps(c, st);
  if (st.mat.normalTextureId >= 0) applyNormalMap(*c.s, st, c.pc->useFaceNormal);
  V3<R> kd = load3<R>(st.mat.diffuse), eta = load3<R>(st.mat.radiance), k = load3<R>(st.mat.radianceFactor);
  directLight(
      c, pl, st, [&](const LRec<R>&) { return V3<R>(0); },
      [&](const LRec<R>& l) {
        R NdotL = dot(l.d, st.ffN), NdotV = dot(st.V, st.ffN);
        if (NdotL < R(0) || NdotV < R(0) || (l.flags & EDelta) == 0) return R(0);
        return std::fabs(dot(reflect(-st.V, st.ffN), l.d) - R(1)) < K<R>::EPS ? R(1) : R(0);
      });
  pl.fromDielectric = true;
  (void)rnd<R>(pl.seed);
  (void)rnd<R>(pl.seed);
  R NdotV = dot(st.V, st.ffN);
  if (NdotV <= R(0)) {
    pl.stop = true;
    return;
  }
  V3<R> bd = reflect(-st.V, st.ffN);
  V3<R> bw = kd * conductorReflectance3(eta, k, dot(st.ffN, bd));
  if (length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  nextRay(pl, st, bd, bw, R(1), ESpecularReflection, true);
}

// brdf_rough_conductor.rchit:72-220
template <class R>
V3<R> roughConductorEval(V3<R> L, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, V3<R> kd, V3<R> eta, V3<R> k, R ax, R ay,
                         uint32_t flags) {
  V3<R> weight(0);
  R NdotL = dot(L, N), NdotV = dot(V, N);
  if ((flags & EArea) == 0 || NdotL < R(0) || NdotV < R(0)) return weight;
  V3<R> H = makeNormal(L + V);
  V3<R> Fs = conductorReflectance3(eta, k, NdotL);
  R Gs = g1SmithAnisoGGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
  Gs *= g1SmithAnisoGGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
  R Ds = dAnisoGGX(dot(H, N), dot(H, X), dot(H, Y), ax, ay);
  return weight + kd * Fs * Gs * Ds * NdotL;
}

template <class R> void chsRoughConductor(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  R t4[4];
  diffuseAndNormalMaps(c, st);
  if (st.mat.roughnessTextureId >= 0) {
    textureEval3(*c.s, st.mat.roughnessTextureId, st.uvx, st.uvy, t4);
    st.mat.anisoAlpha[0] = float(t4[0]), st.mat.anisoAlpha[1] = float(t4[1]);
  }
  if (st.mat.normalTextureId >= 0) applyNormalMap(*c.s, st, c.pc->useFaceNormal);
  V3<R> kd = load3<R>(st.mat.diffuse), eta = load3<R>(st.mat.radiance), k = load3<R>(st.mat.radianceFactor);
  R ax = std::max(R(st.mat.anisoAlpha[0]), K<R>::EPS), ay = std::max(R(st.mat.anisoAlpha[1]), K<R>::EPS);
  directLight(
      c, pl, st,
      [&](const LRec<R>& l) { return roughConductorEval(l.d, st.V, st.ffN, st.X, st.Y, kd, eta, k, ax, ay, EArea); },
      [&](const LRec<R>& l) {
        R NdotL = dot(l.d, st.ffN), NdotV = dot(st.V, st.ffN);
        if (NdotL < R(0) || NdotV < R(0) || (l.flags & EArea) == 0) return R(0);
        V3<R> wi = toLocal(st.X, st.Y, st.ffN, l.d), wo = toLocal(st.X, st.Y, st.ffN, st.V);
        return importanceAnisoGGXPdf(makeNormal(wi + wo), wo, ax, ay);
      });
  R u0 = rnd<R>(pl.seed), u1 = rnd<R>(pl.seed);
  V3<R> N = st.ffN, V = st.V, X = st.X, Y = st.Y;
  R NdotV = dot(V, N);
  if (NdotV <= R(0)) {
    pl.stop = true;
    return;
  }
  V3<R> wo = toLocal(X, Y, N, V);
  V3<R> wi = importanceSampleAnisoGGX(u0, u1, wo, ax, ay);
  V3<R> L = toWorld(X, Y, N, wi);
  V3<R> H = makeNormal(V + L);
  V3<R> wh = makeNormal(wi + wo);
  R NdotL = dot(N, L);
  V3<R> Fs = conductorReflectance3(eta, k, NdotL);
  R Gs = g1SmithAnisoGGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
  Gs *= g1SmithAnisoGGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
  R Ds = dAnisoGGX(dot(H, N), dot(H, X), dot(H, Y), ax, ay);
  R bpdf = importanceAnisoGGXPdf(wh, wo, ax, ay);
  V3<R> bw = kd * Fs * Gs * Ds * NdotL;
  if (bpdf <= R(0) || length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  nextRay(pl, st, L, bw, bpdf, EGlossyReflection, true);
}

// brdf_plastic.rchit:45-93 eval / pdf (smooth coating over a diffuse base).
// The diffuse-lobe pdf is NdotL * (1 - probSpecular), without 1/pi, as written.
template <class R>
V3<R> plasticEval(V3<R> L, V3<R> V, V3<R> N, V3<R> kd, R eta, R fdrInt, R invEta2, uint32_t flags) {
  R NdotL = dot(L, N), NdotV = dot(V, N);
  if (NdotL < R(0) || NdotV < R(0)) return V3<R>(0);
  R Fo = fresnelDielectricExt(NdotV, eta);
  if (flags & EDelta) return std::fabs(dot(reflect(-V, N), L) - R(1)) < K<R>::EPS ? Fo * V3<R>(1) : V3<R>(0);
  if (flags & EArea) {
    R Fi = fresnelDielectricExt(NdotL, eta);
    V3<R> diff = kd / (V3<R>(1) - kd * fdrInt);
    return (R(1) - Fi) * (R(1) - Fo) * diff * invEta2 * K<R>::INV_PI * NdotL;
  }
  return V3<R>(0);
}
template <class R> R plasticPdf(V3<R> L, V3<R> V, V3<R> N, R eta, R ssw, uint32_t flags) {
  R NdotL = dot(L, N), NdotV = dot(V, N);
  if (NdotL < R(0) || NdotV < R(0)) return R(0);
  R Fo = fresnelDielectricExt(NdotV, eta);
  R ps = (Fo * ssw) / (Fo * ssw + (R(1) - Fo) * (R(1) - ssw));
  if (flags & EDelta) return std::fabs(dot(reflect(-V, N), L) - R(1)) < K<R>::EPS ? ps : R(0);
  if (flags & EArea) return NdotL * (R(1) - ps);
  return R(0);
}

// brdf_plastic.rchit:133-204
template <class R> void chsPlastic(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  diffuseAndNormalMaps(c, st);
  if (st.mat.normalTextureId >= 0) applyNormalMap(*c.s, st, c.pc->useFaceNormal);
  V3<R> kd = load3<R>(st.mat.diffuse);
  R eta = R(st.mat.ior), fdrInt = R(st.mat.radiance[0]);
  R dAvg = luminance(kd), sAvg = luminance(V3<R>(1));
  R ssw = sAvg / (dAvg + sAvg);
  R invEta2 = R(1) / (eta * eta);
  V3<R> N = st.ffN, V = st.V;
  auto probSpec = [&](R Fo) { return (Fo * ssw) / (Fo * ssw + (R(1) - Fo) * (R(1) - ssw)); };
  directLight(
      c, pl, st, [&](const LRec<R>& l) { return plasticEval(l.d, V, N, kd, eta, fdrInt, invEta2, EArea); },
      [&](const LRec<R>& l) { return plasticPdf(l.d, V, N, eta, ssw, l.flags); });
  R u0 = rnd<R>(pl.seed), u1 = rnd<R>(pl.seed);
  R NdotV = dot(V, N);
  if (NdotV <= R(0)) {
    pl.stop = true;
    return;
  }
  R Fo = fresnelDielectricExt(NdotV, eta);
  R ps = probSpec(Fo);
  V3<R> bd, bw;
  R bpdf;
  uint32_t bflags;
  if (u0 < ps) {
    bd = makeNormal(reflect(-V, N));
    bflags = ESpecularReflection;
    bpdf = ps;
    bw = V3<R>(1) * Fo;
  } else {
    u0 = (u0 - ps) / (R(1) + K<R>::EPS - ps);
    V3<R> wi = cosineSampleHemisphere(u0, u1);
    R Fi = fresnelDielectricExt(wi.z, eta);
    V3<R> diff = kd / (V3<R>(1) - kd * fdrInt);
    bpdf = (R(1) - ps) * cosineHemispherePdf(wi.z);
    bd = toWorld(st.X, st.Y, N, wi);
    bflags = EDiffuseReflection;
    bw = (R(1) - Fi) * (R(1) - Fo) * invEta2 * diff * K<R>::INV_PI * std::fabs(wi.z);
  }
  if (bpdf <= R(0) || length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  nextRay(pl, st, bd, bw, bpdf, bflags, true);
}

// brdf_phong.rchit:12-193 (shininess in mat.specular, specular albedo in rhoSpec)
template <class R> V3<R> phongEval(V3<R> L, V3<R> V, V3<R> N, V3<R> kd, V3<R> ks, R sh, uint32_t flags) {
  V3<R> weight(0);
  if ((flags & EArea) == 0) return weight;
  R NdotL = dot(N, L), NdotV = dot(N, V);
  if (NdotL < R(0) || NdotV < R(0)) return weight;
  V3<R> H = makeNormal(L + V);
  V3<R> diffuseLobe = kd * K<R>::INV_PI;
  V3<R> specularLobe = ks * std::pow(std::max(dot(H, N), R(0)), sh) * (sh + R(2)) * K<R>::INV_2PI;
  R db = luminance(kd), sb = luminance(ks);
  R dw = db / (db + sb);
  R sw = R(1) - dw;
  return (dw * diffuseLobe + sw * specularLobe) * NdotL;
}
template <class R> R phongLobePdf(V3<R> L, V3<R> V, V3<R> N, R sh) {
  V3<R> H = makeNormal(L + V);
  R NdotH = std::max(dot(H, N), R(0)), VdotH = std::max(dot(H, V), R(0));
  return (sh + R(1)) * K<R>::INV_2PI * std::pow(NdotH, sh) * R(0.25) / (VdotH + K<R>::EPS);
}
template <class R> R phongPdf(V3<R> L, V3<R> V, V3<R> N, V3<R> kd, V3<R> ks, R sh, uint32_t flags) {
  if ((flags & EArea) == 0) return R(0);
  R db = luminance(kd), sb = luminance(ks);
  R dw = db / (db + sb);
  R sw = R(1) - dw;
  return dw * cosineHemispherePdf(dot(N, L)) + sw * phongLobePdf(L, V, N, sh);
}

template <class R> void chsPhong(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  const GpuPushConstantRaytrace& pc = *c.pc;
  diffuseAndNormalMaps(c, st);
  if (st.mat.normalTextureId >= 0) applyNormalMap(*c.s, st, pc.useFaceNormal);
  V3<R> kd = load3<R>(st.mat.diffuse), ks = load3<R>(st.mat.rhoSpec);
  R sh = R(st.mat.specular);
  if (pl.depth == 1) {
    if (pc.diffuseOutChannel >= 0) pl.ch[pc.diffuseOutChannel] = kd;
    if (pc.normalOutChannel >= 0) pl.ch[pc.normalOutChannel] = st.ffN;
    if (pc.specularOutChannel >= 0) pl.ch[pc.specularOutChannel] = ks;
    if (pc.tangentOutChannel >= 0) pl.ch[pc.tangentOutChannel] = st.X;
    if (pc.roughnessOutChannel >= 0) pl.ch[pc.roughnessOutChannel] = V3<R>(1, 1, 0);
    if (pc.positionOutChannel >= 0) pl.ch[pc.positionOutChannel] = st.pos;
    if (pc.uvOutChannel >= 0) pl.ch[pc.uvOutChannel] = V3<R>(st.uvx, st.uvy, 1);
  }
  directLight(
      c, pl, st, [&](const LRec<R>& l) { return phongEval(l.d, st.V, st.ffN, kd, ks, sh, EArea); },
      [&](const LRec<R>& l) { return phongPdf(l.d, st.V, st.ffN, kd, ks, sh, l.flags); });
  R u0 = rnd<R>(pl.seed), u1 = rnd<R>(pl.seed);
  R db = luminance(kd), sb = luminance(ks);
  R dw = db / (db + sb);
  R sw = R(1) - dw;
  V3<R> bd;
  R bpdf;
  uint32_t bflags;
  if (u0 < dw) {
    u0 /= dw;
    V3<R> wi = cosineSampleHemisphere(u0, u1);
    bpdf = cosineHemispherePdf(wi.z);
    bd = toWorld(st.X, st.Y, st.ffN, wi);
    bflags = EDiffuseReflection;
  } else {
    u0 = (u0 - dw) / sw;
    R cosTheta = std::pow(u0, R(1) / (sh + R(1)));
    R phi = K<R>::TWO_PI * u1;
    R sinTheta = safeSqrt(R(1) - cosTheta * cosTheta);
    V3<R> wh(sinTheta * std::sin(phi), sinTheta * std::cos(phi), cosTheta);
    V3<R> H = toWorld(st.X, st.Y, st.ffN, wh);
    V3<R> L = reflect(-st.V, H);
    bpdf = phongLobePdf(L, st.V, st.ffN, sh);
    bd = L;
    bflags = EGlossyReflection;
  }
  V3<R> bw = phongEval(bd, st.V, st.ffN, kd, ks, sh, EArea);
  if (bpdf <= R(0) || length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  nextRay(pl, st, bd, bw, bpdf, bflags, false);
}

// brdf_kang18.rchit:57-249 (GGX + Lambert with its own luminance weights)
template <class R> R lum18(V3<R> c) { return R(0.212671) * c.x + R(0.715160) * c.y + R(0.072169) * c.z; }
template <class R>
V3<R> kangEval(V3<R> L, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, V3<R> kd, V3<R> ks, R ax, R ay, R eta) {
  V3<R> weight(0);
  R NdotL = dot(N, L), NdotV = dot(N, V);
  if (NdotL <= R(0) || NdotV <= R(0)) return weight;
  V3<R> H = makeNormal(L + V);
  R HdotN = dot(H, N), HdotX = dot(H, X), HdotY = dot(H, Y), HdotV = dot(H, V);
  R F0 = sqr((eta - R(1)) / (eta + R(1)));
  R Fs = F0 + (R(1) - F0) * std::pow(std::min(std::max(R(1) - HdotV, R(0)), R(1)), R(5));
  R Ds = dAnisoGGX(HdotN, HdotX, HdotY, ax, ay);
  R Gs = g1SmithAnisoGGX(NdotV, dot(V, X), dot(V, Y), ax, ay);
  Gs *= g1SmithAnisoGGX(NdotL, dot(L, X), dot(L, Y), ax, ay);
  V3<R> diffuse = kd * K<R>::INV_PI;
  V3<R> specular = ks * Fs * Ds * Gs;
  weight = weight + diffuse;
  weight = weight + specular;
  return weight * NdotL;
}
template <class R>
R kangPdf(V3<R> L, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, R ax, R ay, V3<R> kd, V3<R> ks, uint32_t flags) {
  if ((flags & EArea) == 0) return R(0);
  R NdotL = dot(N, L), NdotV = dot(N, V);
  if (NdotL <= R(0) || NdotV <= R(0)) return R(0);
  V3<R> H = makeNormal(L + V);
  V3<R> wh = makeNormal(toLocal(X, Y, N, H)), wo = makeNormal(toLocal(X, Y, N, V)), wi = makeNormal(toLocal(X, Y, N, L));
  R pd = lum18(kd), ps = lum18(ks);
  R P = pd / (pd + ps + K<R>::EPS);
  R pdf = 0;
  pdf += P * cosineHemispherePdf(wi.z);
  pdf += (R(1) - P) * importanceAnisoGGXPdf(wh, wo, ax, ay);
  return pdf;
}

template <class R> void chsKang18(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  const orc_scene& s = *c.s;
  const GpuPushConstantRaytrace& pc = *c.pc;
  R t4[4];
  diffuseAndNormalMaps(c, st);
  if (st.mat.metalnessTextureId >= 0) {
    V3<R> v = textureEval3(s, st.mat.metalnessTextureId, st.uvx, st.uvy, t4);
    st.mat.rhoSpec[0] = float(v.x), st.mat.rhoSpec[1] = float(v.y), st.mat.rhoSpec[2] = float(v.z);
  }
  if (st.mat.roughnessTextureId >= 0) {
    textureEval3(s, st.mat.roughnessTextureId, st.uvx, st.uvy, t4);
    st.mat.anisoAlpha[0] = float(t4[0]), st.mat.anisoAlpha[1] = float(t4[1]);
  }
  R opacity = R(st.mat.metalness);
  if (st.mat.opacityTextureId >= 0) {
    textureEval3(s, st.mat.opacityTextureId, st.uvx, st.uvy, t4);
    opacity = t4[0];
  }
  const float* w = st.w2o;
  auto objToWorldNormal = [&](V3<R> n) {  // (n * gl_WorldToObjectEXT).xyz
    return V3<R>(n.x * R(w[0]) + n.y * R(w[4]) + n.z * R(w[8]), n.x * R(w[1]) + n.y * R(w[5]) + n.z * R(w[9]),
                 n.x * R(w[2]) + n.y * R(w[6]) + n.z * R(w[10]));
  };
  if (st.mat.normalTextureId >= 0) {
    V3<R> cn = textureEval3(s, st.mat.normalTextureId, st.uvx, st.uvy, t4);
    st.N = makeNormal(objToWorldNormal(R(2) * cn - V3<R>(1)));
    if (pc.useFaceNormal == 1) st.N = V3<R>(0);
    basis(st.N, st.X, st.Y);
    st.ffN = dot(st.N, st.V) > R(0) ? st.N : -st.N;
  }
  if (st.mat.tangentTextureId >= 0) {
    V3<R> ct = textureEval3(s, st.mat.tangentTextureId, st.uvx, st.uvy, t4);
    st.X = makeNormal(objToWorldNormal(R(2) * ct - V3<R>(1)));
  }
  st.Y = makeNormal(cross(st.N, st.X));
  st.X = makeNormal(cross(st.Y, st.N));
  st.ffN = dot(st.N, st.V) > R(0) ? st.N : -st.N;
  if (rnd<R>(pl.seed) < opacity) {
    pl.ro = offsetPositionAlongNormal(st.pos, -st.ffN);
    pl.depth--;
    return;
  }
  R ax = std::max(R(st.mat.anisoAlpha[0]), K<R>::EPS), ay = std::max(R(st.mat.anisoAlpha[1]), K<R>::EPS);
  R eta = R(st.mat.ior);
  V3<R> kd = load3<R>(st.mat.diffuse), ks = load3<R>(st.mat.rhoSpec);
  if (pl.depth == 1) {
    if (pc.diffuseOutChannel >= 0) pl.ch[pc.diffuseOutChannel] = kd;
    if (pc.normalOutChannel >= 0) pl.ch[pc.normalOutChannel] = st.ffN;
    if (pc.specularOutChannel >= 0) pl.ch[pc.specularOutChannel] = ks;
    if (pc.tangentOutChannel >= 0) pl.ch[pc.tangentOutChannel] = st.X;
    if (pc.roughnessOutChannel >= 0) pl.ch[pc.roughnessOutChannel] = V3<R>(ax, ay, 0);
    if (pc.positionOutChannel >= 0) pl.ch[pc.positionOutChannel] = st.pos;
    if (pc.uvOutChannel >= 0) pl.ch[pc.uvOutChannel] = V3<R>(st.uvx, st.uvy, 1);
  }
  directLight(
      c, pl, st, [&](const LRec<R>& l) { return kangEval(l.d, st.V, st.ffN, st.X, st.Y, kd, ks, ax, ay, eta); },
      [&](const LRec<R>& l) { return kangPdf(l.d, st.V, st.ffN, st.X, st.Y, ax, ay, kd, ks, l.flags); });
  R u0 = rnd<R>(pl.seed), u1 = rnd<R>(pl.seed);
  V3<R> wo = makeNormal(toLocal(st.X, st.Y, st.ffN, st.V));
  R pd = lum18(kd), ps = lum18(ks);
  R P = pd / (pd + ps + K<R>::EPS);
  V3<R> wi;
  R bpdf;
  uint32_t bflags;
  if (rnd<R>(pl.seed) < P) {
    wi = cosineSampleHemisphere(u0, u1);
    bflags = EDiffuseReflection;
    bpdf = cosineHemispherePdf(wi.z);
  } else {
    wi = importanceSampleAnisoGGX(u0, u1, wo, ax, ay);
    bpdf = importanceAnisoGGXPdf(makeNormal(wi + wo), wo, ax, ay);
    bflags = EGlossyReflection;
  }
  V3<R> bd = toWorld(st.X, st.Y, st.ffN, wi);
  V3<R> bw = kangEval(bd, st.V, st.ffN, st.X, st.Y, kd, ks, ax, ay, eta);
  if (bpdf <= R(0) || length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  nextRay(pl, st, bd, bw, bpdf, bflags, true);
}

// brdf_disney.rchit:31-405 (the reflection lobes: diffuse with retro-
// reflection / fake subsurface / sheen, anisotropic GGX through visible-
// normal sampling, GTR1 clearcoat).  GLSL mix(x, y, a) = x (1 - a) + y a.
template <class R> struct DisneyMat {
  V3<R> baseColor;
  R anisotropic, metallic, roughness, subsurface, specularTint, sheen, sheenTint, clearcoat, clearcoatRoughness, ior,
      ax, ay;
};
template <class R> R mixf(R x, R y, R a) { return x * (R(1) - a) + y * a; }
template <class R> R clamp01(R x) { return std::fmin(std::fmax(x, R(0)), R(1)); }
template <class R> R dsnLum(V3<R> c) { return R(0.212671) * c.x + R(0.715160) * c.y + R(0.072169) * c.z; }
template <class R> R dsnGTR1(R NDotH, R a) {
  if (a >= R(1)) return K<R>::INV_PI;
  R a2 = a * a;
  R t = R(1) + (a2 - R(1)) * NDotH * NDotH;
  return (a2 - R(1)) / (K<R>::PI * std::log(a2) * t);
}
template <class R> V3<R> dsnSampleGTR1(R rgh, R r1, R r2) {
  R a = std::max(R(0.001), rgh);
  R a2 = a * a;
  R phi = r1 * K<R>::TWO_PI;
  R cosTheta = std::sqrt((R(1) - std::pow(a2, R(1) - r1)) / (R(1) - a2));
  R sinTheta = clamp01(std::sqrt(R(1) - (cosTheta * cosTheta)));
  return V3<R>(sinTheta * std::cos(phi), sinTheta * std::sin(phi), cosTheta);
}
template <class R> V3<R> dsnSampleVNDF(V3<R> V, R ax, R ay, R r1, R r2) {
  V3<R> Vh = makeNormal(V3<R>(ax * V.x, ay * V.y, V.z));
  R lensq = Vh.x * Vh.x + Vh.y * Vh.y;
  V3<R> T1 = lensq > R(0) ? V3<R>(-Vh.y, Vh.x, R(0)) * (R(1) / std::sqrt(lensq)) : V3<R>(1, 0, 0);
  V3<R> T2 = cross(Vh, T1);
  R r = std::sqrt(r1);
  R phi = R(2) * K<R>::PI * r2;
  R t1 = r * std::cos(phi), t2 = r * std::sin(phi);
  R sv = R(0.5) * (R(1) + Vh.z);
  t2 = (R(1) - sv) * std::sqrt(R(1) - t1 * t1) + sv * t2;
  V3<R> Nh = t1 * T1 + t2 * T2 + std::sqrt(std::max(R(0), R(1) - t1 * t1 - t2 * t2)) * Vh;
  return makeNormal(V3<R>(ax * Nh.x, ay * Nh.y, std::max(R(0), Nh.z)));
}
template <class R> R dsnGTR2Aniso(R NDotH, R HDotX, R HDotY, R ax, R ay) {
  R a = HDotX / ax, b = HDotY / ay;
  R c = a * a + b * b + NDotH * NDotH;
  return R(1) / (K<R>::PI * ax * ay * c * c);
}
template <class R> R dsnSmithG(R NDotV, R alphaG) {
  R a = alphaG * alphaG, b = NDotV * NDotV;
  return (R(2) * NDotV) / (NDotV + std::sqrt(a + b - a * b));
}
template <class R> R dsnSmithGAniso(R NDotV, R VDotX, R VDotY, R ax, R ay) {
  R a = VDotX * ax, b = VDotY * ay, c = NDotV;
  return (R(2) * NDotV) / (NDotV + std::sqrt(a * a + b * b + c * c));
}
template <class R> R dsnSchlick(R u) {
  R m = clamp01(R(1) - u);
  R m2 = m * m;
  return m2 * m2 * m;
}
template <class R> R dsnDielectricFresnel(R cosThetaI, R eta) {
  R sinThetaTSq = eta * eta * (R(1) - cosThetaI * cosThetaI);
  if (sinThetaTSq > R(1)) return R(1);
  R cosThetaT = std::sqrt(std::max(R(1) - sinThetaTSq, R(0)));
  R rs = (eta * cosThetaT - cosThetaI) / (eta * cosThetaT + cosThetaI);
  R rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
  return R(0.5) * (rs * rs + rp * rp);
}
template <class R> R dsnFresnel(R metallic, R eta, R LDotH, R VDotH) {
  return mixf(dsnDielectricFresnel(std::fabs(VDotH), eta), dsnSchlick(LDotH), metallic);
}
template <class R> V3<R> dsnEvalDiffuse(const DisneyMat<R>& m, V3<R> Csheen, V3<R> V, V3<R> L, V3<R> H, R& pdf) {
  pdf = 0;
  if (L.z <= R(0)) return V3<R>(0);
  R FL = dsnSchlick(L.z), FV = dsnSchlick(V.z), FH = dsnSchlick(dot(L, H));
  R Fd90 = R(0.5) + R(2) * dot(L, H) * dot(L, H) * m.roughness;
  R Fd = mixf(R(1), Fd90, FL) * mixf(R(1), Fd90, FV);
  R Fss90 = dot(L, H) * dot(L, H) * m.roughness;
  R Fss = mixf(R(1), Fss90, FL) * mixf(R(1), Fss90, FV);
  R ss = R(1.25) * (Fss * (R(1) / (L.z + V.z) - R(0.5)) + R(0.5));
  V3<R> Fsheen = FH * m.sheen * Csheen;
  pdf = L.z * K<R>::INV_PI;
  return (R(1) - m.metallic) * (K<R>::INV_PI * mixf(Fd, ss, m.subsurface) * m.baseColor + Fsheen);
}
template <class R>
V3<R> dsnEvalSpec(const DisneyMat<R>& m, R eta, V3<R> specCol, V3<R> V, V3<R> L, V3<R> H, R& pdf) {
  pdf = 0;
  if (L.z <= R(0)) return V3<R>(0);
  R FM = dsnFresnel(m.metallic, eta, dot(L, H), dot(V, H));
  V3<R> F = mixv(specCol, V3<R>(1), FM);
  R D = dsnGTR2Aniso(H.z, H.x, H.y, m.ax, m.ay);
  R G1 = dsnSmithGAniso(std::fabs(V.z), V.x, V.y, m.ax, m.ay);
  R G2 = G1 * dsnSmithGAniso(std::fabs(L.z), L.x, L.y, m.ax, m.ay);
  pdf = G1 * D / (R(4) * V.z);
  return F * D * G2 / (R(4) * L.z * V.z);
}
template <class R> V3<R> dsnEvalClearcoat(const DisneyMat<R>& m, V3<R> V, V3<R> L, V3<R> H, R& pdf) {
  pdf = 0;
  if (L.z <= R(0)) return V3<R>(0);
  R FH = dsnDielectricFresnel(dot(V, H), R(1) / R(1.5));
  R F = mixf(R(0.04), R(1), FH);
  R D = dsnGTR1(H.z, m.clearcoatRoughness);
  R G = dsnSmithG(L.z, R(0.25)) * dsnSmithG(V.z, R(0.25));
  R jacobian = R(1) / (R(4) * dot(V, H));
  pdf = D * H.z * jacobian;
  return V3<R>(0.25) * m.clearcoat * F * D * G / (R(4) * L.z * V.z);
}
template <class R> void dsnSpecColor(const DisneyMat<R>& m, R eta, V3<R>& specCol, V3<R>& sheenCol) {
  R lum = dsnLum(m.baseColor);
  V3<R> ctint = lum > R(0) ? m.baseColor / lum : V3<R>(1);
  R F0 = (R(1) - eta) / (R(1) + eta);
  specCol = mixv(F0 * F0 * mixv(V3<R>(1), ctint, m.specularTint), m.baseColor, m.metallic);
  sheenCol = mixv(V3<R>(1), ctint, m.sheenTint);
}
template <class R> void dsnLobes(const DisneyMat<R>& m, V3<R> specCol, R approxFresnel, R& dw, R& sw, R& cw) {
  dw = dsnLum(m.baseColor) * (R(1) - m.metallic);
  sw = dsnLum(mixv(specCol, V3<R>(1), approxFresnel));
  cw = R(0.25) * m.clearcoat * (R(1) - m.metallic);
  R total = dw + sw + cw;
  dw /= total;
  sw /= total;
  cw /= total;
}
template <class R>
V3<R> dsnEval(V3<R> L, V3<R> V, V3<R> N, V3<R> X, V3<R> Y, const DisneyMat<R>& m, R eta, uint32_t flags, R& bsdfPdf) {
  V3<R> weight(0);
  bsdfPdf = 0;
  if ((flags & EArea) == 0) return weight;
  V = toLocal(X, Y, N, V);
  L = toLocal(X, Y, N, L);
  if (L.z <= R(0) || V.z <= R(0)) return weight;
  V3<R> H = makeNormal(L + V);
  if (H.z < R(0)) H = -H;
  V3<R> specCol, sheenCol;
  dsnSpecColor(m, eta, specCol, sheenCol);
  R dw, sw, cw;
  dsnLobes(m, specCol, dsnFresnel(m.metallic, eta, dot(L, H), dot(V, H)), dw, sw, cw);
  R pdf = 0;
  if (dw > R(0) && L.z > R(0)) {
    weight = weight + dsnEvalDiffuse(m, sheenCol, V, L, H, pdf);
    bsdfPdf += pdf * dw;
  }
  if (sw > R(0) && L.z > R(0) && V.z > R(0)) {
    weight = weight + dsnEvalSpec(m, eta, specCol, V, L, H, pdf);
    bsdfPdf += pdf * sw;
  }
  if (cw > R(0) && L.z > R(0) && V.z > R(0)) {
    weight = weight + dsnEvalClearcoat(m, V, L, H, pdf);
    bsdfPdf += pdf * cw;
  }
  return weight * L.z;
}
// sampleBsdf, brdf_disney.rchit:298-363.  The returned weight is
// f * |dot(N, L)| with N in world space and L still local, as written.
template <class R>
V3<R> dsnSample(R r1, R r2, V3<R> Vw, V3<R> N, V3<R> X, V3<R> Y, const DisneyMat<R>& m, R eta, V3<R>& d, R& pdf,
                uint32_t& flags) {
  pdf = 0;
  V3<R> f(0);
  V3<R> V = toLocal(X, Y, N, Vw), L;
  V3<R> specCol, sheenCol;
  dsnSpecColor(m, eta, specCol, sheenCol);
  R dw, sw, cw;
  dsnLobes(m, specCol, dsnFresnel(m.metallic, eta, V.z, V.z), dw, sw, cw);
  R cdf0 = dw, cdf1 = cdf0 + cw;
  if (r1 < cdf0) {
    r1 /= cdf0;
    L = cosineSampleHemisphere(r1, r2);
    V3<R> H = makeNormal(L + V);
    f = dsnEvalDiffuse(m, sheenCol, V, L, H, pdf);
    pdf *= dw;
    flags = EDiffuseReflection;
  } else if (r1 < cdf1) {
    r1 = (r1 - cdf0) / (cdf1 - cdf0);
    V3<R> H = dsnSampleGTR1(m.clearcoatRoughness, r1, r2);
    if (H.z < R(0)) H = -H;
    L = makeNormal(reflect(-V, H));
    f = dsnEvalClearcoat(m, V, L, H, pdf);
    pdf *= cw;
    flags = EGlossyReflection;
  } else {
    r1 = (r1 - cdf1) / (R(1) - cdf1);
    V3<R> H = dsnSampleVNDF(V, m.ax, m.ay, r1, r2);
    if (H.z < R(0)) H = -H;
    L = makeNormal(reflect(-V, H));
    f = dsnEvalSpec(m, eta, specCol, V, L, H, pdf);
    pdf *= sw;
    flags = EGlossyReflection;
  }
  d = toWorld(X, Y, N, L);
  return f * std::fabs(dot(N, L));
}
template <class R> DisneyMat<R> disneyMat(const GpuMaterial& g, R rough) {
  DisneyMat<R> m;
  R aspect = std::sqrt(R(1) - R(g.anisotropic) * R(0.9));
  m.ax = std::max(R(0.001), rough * rough / aspect);
  m.ay = std::max(R(0.001), rough * rough * aspect);
  m.baseColor = load3<R>(g.diffuse);
  m.anisotropic = R(g.anisotropic);
  m.metallic = R(g.metalness);
  m.roughness = std::max(rough * rough, R(0.001));
  m.subsurface = R(g.subsurface);
  m.specularTint = R(g.specularTint);
  m.sheen = R(g.sheen);
  m.sheenTint = R(g.sheenTint);
  m.clearcoat = R(g.clearcoat);
  m.clearcoatRoughness = mixf(R(0.1), R(0.001), R(g.clearcoatGloss));
  m.ior = R(g.ior);
  return m;
}

template <class R> void chsDisney(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  const orc_scene& s = *c.s;
  R t4[4];
  diffuseAndNormalMaps(c, st);
  if (st.mat.metalnessTextureId >= 0) {
    textureEval3(s, st.mat.metalnessTextureId, st.uvx, st.uvy, t4);
    st.mat.metalness = float(t4[0]);
  }
  if (st.mat.roughnessTextureId >= 0) {
    textureEval3(s, st.mat.roughnessTextureId, st.uvx, st.uvy, t4);
    st.mat.roughness = float(t4[0]);
  }
  if (st.mat.normalTextureId >= 0) applyNormalMap(s, st, c.pc->useFaceNormal);
  R opacity = R(st.mat.rhoSpec[0]);
  if (st.mat.opacityTextureId >= 0) {
    textureEval3(s, st.mat.opacityTextureId, st.uvx, st.uvy, t4);
    opacity = t4[0];
  }
  if (rnd<R>(pl.seed) < opacity) {
    pl.ro = offsetPositionAlongNormal(st.pos, -st.ffN);
    pl.depth--;
    return;
  }
  DisneyMat<R> m = disneyMat<R>(st.mat, R(st.mat.roughness));
  R eta = dot(st.V, st.N) > R(0) ? (R(1) / m.ior) : m.ior;
  {
    V3<R> Ld(0);
    bool visible;
    LRec<R> l;
    V3<R> radiance = sampleLights(c, pl, st.pos, st.ffN, visible, l);
    if (visible) {
      R bsdfPdf;
      V3<R> bw = dsnEval(l.d, st.V, st.ffN, st.X, st.Y, m, eta, EArea, bsdfPdf);
      R mis = powerHeuristic(l.pdf, bsdfPdf);
      Ld = Ld + mis * bw * radiance * pl.throughput / (l.pdf + K<R>::EPS);
    }
    pl.dRad = Ld;
    pl.dSkip = !visible;
  }
  R u0 = rnd<R>(pl.seed), u1 = rnd<R>(pl.seed);
  V3<R> bd;
  R bpdf;
  uint32_t bflags = EBsdfNull;
  V3<R> bw = dsnSample(u0, u1, st.V, st.ffN, st.X, st.Y, m, eta, bd, bpdf, bflags);
  if (bpdf <= R(0) || length(bw) == R(0)) {
    pl.stop = true;
    return;
  }
  nextRay(pl, st, bd, bw, bpdf, bflags, false);
}

template <class R> void chsEmissive(const RCtx& c, Payload<R>& pl, HitState<R>& st) {
  // brdf_emissive.rchit:12-26
  pl.stop = true;
  V3<R> rad = load3<R>(st.mat.radiance);
  if (st.mat.radianceTextureId >= 0) {
    R t4[4];
    V3<R> tv = textureEval3(*c.s, st.mat.radianceTextureId, st.uvx, st.uvy, t4);
    rad = load3<R>(st.mat.radianceFactor) * tv;
  }
  if (c.pc->ignoreEmissive == 0) pl.radiance = pl.radiance + rad * pl.throughput;
}

template <class R> void missShader(const RCtx& c, Payload<R>& pl) {
  // raytrace.default.rmiss:24-55
  const GpuPushConstantRaytrace& pc = *c.pc;
  pl.stop = true;
  V3<R> env, d = pl.rd;
  const bool sky = c.s->sky.in_use == 1;
  if (sky)
    env = sunAndSky(c.s->sky, d);
  else if (pc.hasEnvMap == 1)
    env = evalEnvmap(*c.s, c.cam->envTransform, R(pc.envMapIntensity), d);
  else
    env = load3<R>(pc.bgColor);
  R mis = 1;
  if (pl.depth != 1 && isNonSpecular(pl.bflags)) {
    R envPdf;
    if (sky)
      envPdf = K<R>::INV_4PI;
    else if (pc.hasEnvMap == 1)
      envPdf = pdfEnvmap(*c.s, c.cam->envTransform, R(pc.envMapResolution[0]), R(pc.envMapResolution[1]), d);
    else
      envPdf = K<R>::INV_4PI;
    mis = powerHeuristic(pl.bpdf, envPdf);
  }
  pl.radiance = pl.radiance + pl.throughput * env * mis;
}

struct Counters {
  uint64_t paths = 0, ext = 0, shadow = 0, dielShadow = 0;
};

// One sample-frame of one pixel: raytrace.projective.rgen:26-179.
// Camera ray at raster position (px, py), raytrace.projective.rgen:30-87:
// perspective (with the thin lens when aperture > 0, lens sample (l0, l1))
// or OpenCV intrinsics.  The lens direction pFocus - pLens is not
// normalised, as in the shader.
template <class R>
void cameraRayAt(const GpuCamera& cam, R px, R py, R l0, R l1, V3<R>& ro, V3<R>& rd) {
  V3<R> origin = transformPoint(cam.cameraToWorld, V3<R>(0));
  if (cam.type == CameraTypePerspective) {
    ro = origin;
    V3<R> pCamera = transformPoint(cam.rasterToCamera, V3<R>(px, py, R(0)));
    V3<R> r = makeNormal(pCamera);
    if (cam.aperture > 0.f) {
      R dx, dy;
      concentricSampleDisk(l0, l1, dx, dy);
      V3<R> pLens(R(cam.aperture) * dx, R(cam.aperture) * dy, R(0));
      R ft = R(cam.focalDistance) / r.z;
      V3<R> pFocus = ft * r;
      ro = transformPoint(cam.cameraToWorld, pLens);
      r = pFocus - pLens;
    }
    rd = transformDirection(cam.cameraToWorld, r);
  } else {
    ro = origin;
    R rx = (px - R(cam.fxfycxcy[2])) / R(cam.fxfycxcy[0]);
    R ry = (py - R(cam.fxfycxcy[3])) / R(cam.fxfycxcy[1]);
    rd = transformDirection(cam.cameraToWorld, V3<R>(rx, ry, R(1)));
  }
}

// The camera sample of pixel (x, y) in frame `frame` with the seed's draws in
// the shader's order: jitter (rgen:41, none on frame 0), then the lens
// (rgen:57, only for a perspective camera with aperture > 0).
template <class R>
void cameraSample(const GpuCamera& cam, uint32_t x, uint32_t y, int32_t frame, uint32_t& seed, R& jx, R& jy,
                  V3<R>& ro, V3<R>& rd) {
  if (frame == 0) {
    jx = jy = R(0.5);
  } else {
    jx = rnd<R>(seed);
    jy = rnd<R>(seed);
  }
  R l0 = 0, l1 = 0;
  if (cam.type == CameraTypePerspective && cam.aperture > 0.f) {
    l0 = rnd<R>(seed);
    l1 = rnd<R>(seed);
  }
  cameraRayAt(cam, R(x) + jx, R(y) + jy, l0, l1, ro, rd);
}

template <class R>
void renderPixel(const RCtx& c, uint32_t x, uint32_t y, int32_t frame, R& wsum, V3<R>& rws, V3<R>* aov,
                 R& firstT, V3<R>& Lout, Counters& cnt) {
  const GpuCamera& cam = *c.cam;
  const GpuPushConstantRaytrace& pc = *c.pc;
  Payload<R> pl;
  pl.seed = xxhash32Seed(x, y, uint32_t(frame));
  R jx, jy;
  V3<R> ro, rd;
  cameraSample(cam, x, y, frame, pl.seed, jx, jy, ro, rd);
  pl.ro = ro;
  pl.rd = rd;
  pl.stop = false;
  pl.radiance = V3<R>(0);
  pl.throughput = V3<R>(1);
  pl.bflags = EBsdfNull;
  for (int i = 0; i < 8; ++i) pl.ch[i] = V3<R>(0);
  firstT = 0;
  for (pl.depth = 1; int(pl.depth) <= pc.maxPathDepth;) {
    pl.dSkip = true;
    pl.dRad = V3<R>(0);
    pl.fromDielectric = false;
    cnt.ext++;
    HitR h = traceClosest<R>(*c.s, pl.ro, pl.rd, K<R>::MINIMUM, K<R>::INF, false);
    if (!h.hit) {
      missShader(c, pl);
    } else {
      if (pl.depth == 1 && firstT == R(0)) firstT = R(h.t);
      HitState<R> st = getHitState<R>(*c.s, h, pl.rd, pc.useFaceNormal);
      if (st.lightId >= 0) chsLambertian(c, pl, st);
      else {
        switch (st.mat.type) {
          case MaterialTypeBrdfLambertian: chsLambertian(c, pl, st); break;
          case MaterialTypeBrdfPbrMetalnessRoughness: chsPbr(c, pl, st); break;
          case MaterialTypeBrdfRoughPlastic: chsRoughPlastic(c, pl, st); break;
          case MaterialTypeBsdfDielectric: chsDielectric(c, pl, st); break;
          case MaterialTypeBrdfEmissive: chsEmissive(c, pl, st); break;
          case MaterialTypeBrdfMirror: chsMirror(c, pl, st); break;
          case MaterialTypeBrdfConductor: chsConductor(c, pl, st); break;
          case MaterialTypeBrdfRoughConductor: chsRoughConductor(c, pl, st); break;
          case MaterialTypeBrdfPlastic: chsPlastic(c, pl, st); break;
          case MaterialTypeBrdfPhong: chsPhong(c, pl, st); break;
          case MaterialTypeBrdfKang18: chsKang18(c, pl, st); break;
          case MaterialTypeBrdfDisney: chsDisney(c, pl, st); break;
          default: pl.stop = true; break;
        }
      }
    }
    if (!pl.dSkip) {
      R maxDist = pl.dDist - R(2) * K<R>::EPS;
      cnt.shadow++;
      if (pl.fromDielectric) cnt.dielShadow++;
      bool occluded = traceAny<R>(*c.s, pl.dO, pl.dD, R(0), maxDist, false);
      if (!occluded) pl.radiance = pl.radiance + pl.dRad;
    }
    if (pl.stop) break;
    pl.depth++;
  }
  // Gaussian filter weight, rgen:134-145; exp(-8) is folded to its
  // correctly rounded value.
  const R alpha = R(-2);
  const R expXY = sizeof(R) == 4 ? R(0.00033546262f) : R(std::exp(-8.0));
  R ox = jx - R(0.5), oy = jy - R(0.5);
  R fw = std::max(R(0), std::exp(alpha * ox * ox) - expXY) * std::max(R(0), std::exp(alpha * oy * oy) - expXY);
  V3<R> L(clampr(pl.radiance.x, R(0), R(10)), clampr(pl.radiance.y, R(0), R(10)),
          clampr(pl.radiance.z, R(0), R(10)));
  rws = fw * L;
  wsum = fw;
  Lout = L;
  for (int i = 0; i < 8; ++i) aov[i] = pl.ch[i];
  cnt.paths++;
}

}  // namespace

// =================================================================== C API
namespace {
template <class F> void parallelRows(uint32_t y0, uint32_t y1, int threads, F f) {
  if (threads <= 1) {
    for (uint32_t y = y0; y < y1; ++y) f(y);
    return;
  }
  std::atomic<uint32_t> next(y0);
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t)
    ts.emplace_back([&]() {
      for (;;) {
        uint32_t y = next.fetch_add(1);
        if (y >= y1) break;
        f(y);
      }
    });
  for (auto& t : ts) t.join();
}

}  // namespace

extern "C" {

uint32_t orc_xxhash32_seed(uint32_t x, uint32_t y, uint32_t z) { return xxhash32Seed(x, y, z); }
uint32_t orc_pcg(uint32_t* state) { return pcg(*state); }
void orc_rand_stream(uint32_t seed, int n, float* out) {
  for (int i = 0; i < n; ++i) out[i] = rnd<float>(seed);
}

// nvmath-style helpers for the camera restatement (column-major float).
static void m4mul(const float* a, const float* b, float* o) {
  float t[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      t[c * 4 + r] = a[0 * 4 + r] * b[c * 4 + 0] + a[1 * 4 + r] * b[c * 4 + 1] + a[2 * 4 + r] * b[c * 4 + 2] +
                     a[3 * 4 + r] * b[c * 4 + 3];
  std::memcpy(o, t, sizeof t);
}
static void m4ident(float* m) {
  std::memset(m, 0, 64);
  m[0] = m[5] = m[10] = m[15] = 1.f;
}
static void m4scale(float* m, float x, float y, float z) {
  m4ident(m);
  m[0] = x, m[5] = y, m[10] = z;
}

void orc_camera_perspective(uint32_t w, uint32_t h, float fov, const float eye[3], const float lookat[3],
                            const float up[3], float aperture, float focal, const float env[16],
                            GpuCamera* out) {
  std::memset(out, 0, sizeof *out);
  // look_at (gluLookAt convention) then getView = scale(1,-1,-1) * view
  // (camera.h:26-28).
  V3<float> e = load3<float>(eye), ctr = load3<float>(lookat), u = load3<float>(up);
  auto nrm = [](V3<float> v) {
    float n = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return n > 0 ? v * (1.f / n) : V3<float>(0);
  };
  V3<float> z = nrm(e - ctr);
  V3<float> x = cross(u, z);
  V3<float> y = cross(z, x);
  x = nrm(x);
  y = nrm(y);
  float view[16];
  m4ident(view);
  view[0] = x.x, view[4] = x.y, view[8] = x.z;
  view[1] = y.x, view[5] = y.y, view[9] = y.z;
  view[2] = z.x, view[6] = z.y, view[10] = z.z;
  view[12] = -(x.x * e.x + x.y * e.y + x.z * e.z);
  view[13] = -(y.x * e.x + y.y * e.y + y.z * e.z);
  view[14] = -(z.x * e.x + z.y * e.y + z.z * e.z);
  float flip[16];
  m4scale(flip, 1.f, -1.f, -1.f);
  float w2c[16];
  m4mul(flip, view, w2c);
  // invert_rot_trans: R^T, -R^T t (computed in double, rounded once)
  float* c2w = out->cameraToWorld;
  m4ident(c2w);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) c2w[c * 4 + r] = w2c[r * 4 + c];
  for (int r = 0; r < 3; ++r) {
    // -(R^T t)_r = -sum_k R(k, r) t_k with R(k, r) stored at [r*4 + k]
    double t = -(double(w2c[r * 4 + 0]) * w2c[12] + double(w2c[r * 4 + 1]) * w2c[13] + double(w2c[r * 4 + 2]) * w2c[14]);
    c2w[12 + r] = float(t);
  }
  // cameraToRasterTransform (camera.cpp:28-62) in double, inverted in double.
  double aspect = double(float(w) / float(h));
  double ctot = 1.0 / std::tan(double(fov) * 3.14159265358979323846 / 180.0 * 0.5);
  double nearz = 0.1, farz = 100.0, recip = 1.0 / (farz - nearz);
  // cameraToRaster = S(W,H,1) S(.5,.5,1) T(1,1,0) S(1,aspect,1) P
  // x_r = W*0.5*(ctot*x/z + 1), y_r = H*0.5*(aspect*ctot*y/z + 1), z_r = (far*recip*z - near*far*recip)/z
  double M[4][4] = {{0}};
  M[0][0] = w * 0.5 * ctot;
  M[0][2] = w * 0.5;
  M[1][1] = h * 0.5 * aspect * ctot;
  M[1][2] = h * 0.5;
  M[2][2] = farz * recip;
  M[2][3] = -nearz * farz * recip;
  M[3][2] = 1.0;
  // Inverse by Gauss-Jordan in double.
  double A[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 8; ++c) A[r][c] = c < 4 ? M[r][c] : (c - 4 == r ? 1.0 : 0.0);
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int r = col + 1; r < 4; ++r)
      if (std::fabs(A[r][col]) > std::fabs(A[piv][col])) piv = r;
    for (int c = 0; c < 8; ++c) std::swap(A[col][c], A[piv][c]);
    double p = A[col][col];
    for (int c = 0; c < 8; ++c) A[col][c] /= p;
    for (int r = 0; r < 4; ++r)
      if (r != col) {
        double f = A[r][col];
        for (int c = 0; c < 8; ++c) A[r][c] -= f * A[col][c];
      }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out->rasterToCamera[c * 4 + r] = float(A[r][c + 4]);
  if (env) std::memcpy(out->envTransform, env, 64);
  else m4ident(out->envTransform);
  out->type = CameraTypePerspective;
  out->aperture = aperture;
  out->focalDistance = focal;
}

void orc_camera_ray_lens(const GpuCamera* cam, float px, float py, float l0, float l1, float o[3], float d[3]) {
  V3<float> ro, rd;
  cameraRayAt(*cam, px, py, l0, l1, ro, rd);
  o[0] = ro.x, o[1] = ro.y, o[2] = ro.z;
  d[0] = rd.x, d[1] = rd.y, d[2] = rd.z;
}

void orc_camera_ray(const GpuCamera* cam, float px, float py, float o[3], float d[3]) {
  V3<float> origin = transformPoint(cam->cameraToWorld, V3<float>(0));
  V3<float> pCamera = transformPoint(cam->rasterToCamera, V3<float>(px, py, 0.f));
  V3<float> r = makeNormal(pCamera);
  V3<float> dir = transformDirection(cam->cameraToWorld, r);
  o[0] = origin.x, o[1] = origin.y, o[2] = origin.z;
  d[0] = dir.x, d[1] = dir.y, d[2] = dir.z;
}

void orc_env_tables(uint32_t w, uint32_t h, const float* rgba, float* marginal, float* conditional) {
  std::vector<float> m, c;
  envTables(w, h, rgba, m, c);
  std::memcpy(marginal, m.data(), m.size() * 4);
  std::memcpy(conditional, c.data(), c.size() * 4);
}

void orc_texture_bilinear(uint32_t w, uint32_t h, const float* rgba, float u, float v, float out[4]) {
  Img im;
  im.w = w;
  im.h = h;
  im.px.assign(rgba, rgba + size_t(w) * h * 4);
  tex2D(im, u, v, out);
}

int orc_scene_create(const ash_scene_desc* d, orc_scene** out) {
  orc_scene* s = new orc_scene;
  s->verts.resize(d->num_meshes);
  s->idx.resize(d->num_meshes);
  for (uint32_t m = 0; m < d->num_meshes; ++m) {
    const ash_mesh_desc& md = d->meshes[m];
    s->verts[m].assign(md.vertices, md.vertices + md.num_vertices);
    s->idx[m].assign(md.indices, md.indices + md.num_indices);
  }
  s->mats.assign(d->materials, d->materials + d->num_materials);
  s->lights.assign(d->lights, d->lights + d->num_lights);
  s->textures.resize(d->num_textures);
  for (uint32_t t = 0; t < d->num_textures; ++t) {
    s->textures[t].w = d->textures[t].width;
    s->textures[t].h = d->textures[t].height;
    s->textures[t].px.assign(d->textures[t].rgba, d->textures[t].rgba + size_t(s->textures[t].w) * s->textures[t].h * 4);
  }
  if (d->envmap.rgba) {
    s->hasEnv = true;
    s->env.w = d->envmap.width;
    s->env.h = d->envmap.height;
    s->env.px.assign(d->envmap.rgba, d->envmap.rgba + size_t(s->env.w) * s->env.h * 4);
    std::vector<float> m, c;
    envTables(s->env.w, s->env.h, d->envmap.rgba, m, c);
    s->marg.w = s->env.w;
    s->marg.h = s->env.h;
    s->marg.col0_only = true;
    s->marg.px.assign(size_t(s->env.h) * 4, 0.f);
    for (uint32_t j = 0; j < s->env.h; ++j) s->marg.px[j * 4 + 0] = m[j * 2], s->marg.px[j * 4 + 1] = m[j * 2 + 1];
    s->cond.w = s->env.w;
    s->cond.h = s->env.h;
    s->cond.px.assign(size_t(s->env.w) * s->env.h * 4, 0.f);
    for (size_t i = 0; i < size_t(s->env.w) * s->env.h; ++i)
      s->cond.px[i * 4 + 0] = c[i * 2], s->cond.px[i * 4 + 1] = c[i * 2 + 1];
  }
  s->inst.resize(d->num_instances);
  size_t ntri = 0;
  for (uint32_t i = 0; i < d->num_instances; ++i) {
    const ash_instance_desc& id = d->instances[i];
    if (id.mesh_id >= d->num_meshes) {
      delete s;
      return ASH_E_INVALID;
    }
    orc_scene::Inst& in = s->inst[i];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 4; ++c) in.o2w[r * 4 + c] = id.transform[c * 4 + r];
    invert3x4(in.o2w, in.w2o);
    in.mesh = id.mesh_id;
    in.material = id.material_id;
    in.light = id.light_id;
    ntri += d->meshes[id.mesh_id].num_indices / 3;
  }
  s->tri.resize(ntri * 9);
  s->triInst.resize(ntri);
  s->triPrim.resize(ntri);
  size_t g = 0;
  for (uint32_t i = 0; i < d->num_instances; ++i) {
    const orc_scene::Inst& in = s->inst[i];
    const ash_mesh_desc& md = d->meshes[in.mesh];
    for (uint32_t p = 0; p < md.num_indices / 3; ++p, ++g) {
      for (int k = 0; k < 3; ++k) xformPoint3x4(in.o2w, md.vertices[md.indices[p * 3 + k]].pos, &s->tri[g * 9 + k * 3]);
      s->triInst[g] = int32_t(i);
      s->triPrim[g] = int32_t(p);
    }
  }
  // BVH
  Build b;
  b.s = s;
  b.cen.resize(ntri * 3);
  b.blo.resize(ntri * 3);
  b.bhi.resize(ntri * 3);
  for (size_t t = 0; t < ntri; ++t)
    for (int k = 0; k < 3; ++k) {
      float a = s->tri[t * 9 + k], bb = s->tri[t * 9 + 3 + k], cc = s->tri[t * 9 + 6 + k];
      b.blo[t * 3 + k] = std::min(a, std::min(bb, cc));
      b.bhi[t * 3 + k] = std::max(a, std::max(bb, cc));
      b.cen[t * 3 + k] = 0.5f * (b.blo[t * 3 + k] + b.bhi[t * 3 + k]);
    }
  s->order.resize(ntri);
  for (size_t t = 0; t < ntri; ++t) s->order[t] = uint32_t(t);
  if (ntri) {
    s->nodes.reserve(ntri / 2 + 2);
    s->nodes.push_back({});
    buildNode(b, 0, 0, uint32_t(ntri), 0);
  }
  *out = s;
  return ASH_OK;
}

void orc_scene_destroy(orc_scene* s) { delete s; }
uint64_t orc_scene_num_triangles(const orc_scene* s) { return s->triInst.size(); }

static void fillHit(const orc_scene* s, const HitR& h, V3<float> d, ash_hit* out) {
  if (!h.hit) {
    out->t = 1e10f;
    out->u = out->v = 0;
    out->inst_id = -1;
    out->prim_id = -1;
    for (int k = 0; k < 3; ++k) out->normal[k] = 0, out->pos[k] = 0;
    return;
  }
  out->t = float(h.t);
  out->u = float(h.u);
  out->v = float(h.v);
  out->inst_id = s->triInst[h.tri];
  out->prim_id = s->triPrim[h.tri];
  HitState<float> st = getHitState<float>(*s, h, d, 0);
  out->normal[0] = st.N.x, out->normal[1] = st.N.y, out->normal[2] = st.N.z;
  out->pos[0] = st.pos.x, out->pos[1] = st.pos.y, out->pos[2] = st.pos.z;
}

int orc_trace_closest(const orc_scene* s, const float o[3], const float d[3], float tmin, float tmax,
                      ash_hit* out, int brute) {
  V3<float> dd = load3<float>(d);
  HitR h = traceClosest<float>(*s, load3<float>(o), dd, tmin, tmax, brute != 0);
  fillHit(s, h, dd, out);
  return h.hit ? 1 : 0;
}

int orc_trace_any(const orc_scene* s, const float o[3], const float d[3], float tmin, float tmax, int brute) {
  return traceAny<float>(*s, load3<float>(o), load3<float>(d), tmin, tmax, brute != 0) ? 1 : 0;
}

int orc_first_hit(const orc_scene* s, const GpuCamera* cam, uint32_t w, uint32_t h, ash_hit* out, int threads) {
  parallelRows(0, h, threads, [&](uint32_t y) {
    for (uint32_t x = 0; x < w; ++x) {
      // frame 0: pixel centre, and the lens draws of a thin-lens camera
      uint32_t seed = xxhash32Seed(x, y, 0u);
      float jx, jy;
      V3<float> oo, dd;
      cameraSample(*cam, x, y, 0, seed, jx, jy, oo, dd);
      HitR hr = traceClosest<float>(*s, oo, dd, K<float>::MINIMUM, K<float>::INF, false);
      fillHit(s, hr, dd, &out[size_t(y) * w + x]);
    }
  });
  return ASH_OK;
}

int orc_render(const orc_scene* s, const GpuCamera* cam, const GpuPushConstantRaytrace* pc, uint32_t w,
               uint32_t h, int32_t first_frame, int32_t n_frames, uint32_t x0, uint32_t y0, uint32_t x1,
               uint32_t y1, int precision, int threads, float* images, float* sq_out, uint64_t* stats,
               int32_t depth_channel) {
  if (x1 > w || y1 > h || x0 >= x1 || y0 >= y1) return ASH_E_INVALID;
  RCtx c{s, cam, pc};
  std::vector<Counters> cnts(static_cast<size_t>(h));
  size_t plane = size_t(w) * h * 4;
  parallelRows(y0, y1, threads, [&](uint32_t y) {
    Counters& cnt = cnts[y];
    for (uint32_t x = x0; x < x1; ++x) {
      size_t pi = (size_t(y) * w + x) * 4;
      // FP64: keep the running (sum w*L, sum w) in double across the frames
      // of this call, then store the mean.
      double accW = images[8 * plane + pi];
      double accR[3] = {double(images[pi]) * accW, double(images[pi + 1]) * accW, double(images[pi + 2]) * accW};
      for (int32_t f = first_frame; f < first_frame + n_frames; ++f) {
        if (precision == 0) {
          float ws, ft;
          V3<float> rws, aov[8], L;
          renderPixel<float>(c, x, y, f, ws, rws, aov, ft, L, cnt);
          if (f == 0) {
            V3<float> rad = rws / ws;
            images[pi] = rad.x, images[pi + 1] = rad.y, images[pi + 2] = rad.z, images[pi + 3] = 1.f;
            for (uint32_t cid = 0; cid < pc->nMultiChannel; ++cid) {
              float* im = &images[(cid + 1) * plane + pi];
              V3<float> v = aov[cid];
              if (int32_t(cid) == depth_channel) v = V3<float>(ft);
              im[0] = v.x, im[1] = v.y, im[2] = v.z, im[3] = 1.f;
            }
            for (int k = 0; k < 4; ++k) images[8 * plane + pi + k] = ws;
          } else {
            V3<float> oldRad(images[pi], images[pi + 1], images[pi + 2]);
            float oldW = images[8 * plane + pi];
            V3<float> oldRws = oldRad * oldW;
            float newW = oldW + ws;
            V3<float> newRws = oldRws + rws;
            V3<float> rad = newRws / newW;
            images[pi] = rad.x, images[pi + 1] = rad.y, images[pi + 2] = rad.z, images[pi + 3] = 1.f;
            for (int k = 0; k < 4; ++k) images[8 * plane + pi + k] = newW;
          }
          if (sq_out) {
            float* q = &sq_out[pi];
            q[0] += L.x * L.x, q[1] += L.y * L.y, q[2] += L.z * L.z, q[3] += 1.f;
          }
        } else {
          double ws, ft;
          V3<double> rws, aov[8], L;
          renderPixel<double>(c, x, y, f, ws, rws, aov, ft, L, cnt);
          if (f == 0) {
            accW = 0;
            accR[0] = accR[1] = accR[2] = 0;
            for (uint32_t cid = 0; cid < pc->nMultiChannel; ++cid) {
              float* im = &images[(cid + 1) * plane + pi];
              V3<double> v = aov[cid];
              if (int32_t(cid) == depth_channel) v = V3<double>(ft);
              im[0] = float(v.x), im[1] = float(v.y), im[2] = float(v.z), im[3] = 1.f;
            }
          }
          accW += ws;
          accR[0] += rws.x, accR[1] += rws.y, accR[2] += rws.z;
          if (sq_out) {
            float* q = &sq_out[pi];
            q[0] += float(L.x * L.x), q[1] += float(L.y * L.y), q[2] += float(L.z * L.z), q[3] += 1.f;
          }
        }
      }
      if (precision != 0) {
        images[pi] = float(accR[0] / accW), images[pi + 1] = float(accR[1] / accW), images[pi + 2] = float(accR[2] / accW);
        images[pi + 3] = 1.f;
        for (int k = 0; k < 4; ++k) images[8 * plane + pi + k] = float(accW);
      }
    }
  });
  if (stats) {
    Counters t;
    for (auto& c2 : cnts) t.paths += c2.paths, t.ext += c2.ext, t.shadow += c2.shadow, t.dielShadow += c2.dielShadow;
    stats[0] = t.paths, stats[1] = t.ext, stats[2] = t.shadow, stats[3] = t.dielShadow;
  }
  return ASH_OK;
}

// --------------------------------------------------------------- probes
void orc_concentric_disk(float u0, float u1, float out[2]) { concentricSampleDisk(u0, u1, out[0], out[1]); }
void orc_cosine_hemisphere(float u0, float u1, float out[3]) {
  V3<float> v = cosineSampleHemisphere(u0, u1);
  out[0] = v.x, out[1] = v.y, out[2] = v.z;
}
void orc_offset_ray(const float p[3], const float n[3], float out[3]) {
  V3<float> v = offsetPositionAlongNormal(load3<float>(p), load3<float>(n));
  out[0] = v.x, out[1] = v.y, out[2] = v.z;
}
void orc_basis(const float n[3], float f[3], float r[3]) {
  V3<float> a, b;
  basis(load3<float>(n), a, b);
  f[0] = a.x, f[1] = a.y, f[2] = a.z;
  r[0] = b.x, r[1] = b.y, r[2] = b.z;
}
float orc_power_heuristic(float a, float b) { return powerHeuristic(a, b); }

// loader/material.cpp:6-37 (computeDiffuseFresnel via dielectricReflectance).
static float dielectricReflectanceL(float eta, float cosThetaI, float& cosThetaT) {
  if (cosThetaI < 0.0f) {
    eta = 1.0f / eta;
    cosThetaI = -cosThetaI;
  }
  float sinThetaTSq = eta * eta * (1.0f - cosThetaI * cosThetaI);
  if (sinThetaTSq > 1.0f) {
    cosThetaT = 0.0f;
    return 1.0f;
  }
  cosThetaT = std::sqrt(std::max(1.0f - sinThetaTSq, 0.0f));
  float Rs = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
  float Rp = (eta * cosThetaT - cosThetaI) / (eta * cosThetaT + cosThetaI);
  return (Rs * Rs + Rp * Rp) * 0.5f;
}
float orc_compute_diffuse_fresnel(float ior, int sampleCount) {
  double diffuseFresnel = 0.0;
  float t;
  float fb = dielectricReflectanceL(ior, 0.0f, t);
  for (int i = 1; i <= sampleCount; ++i) {
    float cosThetaSq = float(i) / float(sampleCount);
    float fa = dielectricReflectanceL(ior, std::min(std::sqrt(cosThetaSq), 1.0f), t);
    diffuseFresnel += double(fa + fb) * (0.5 / sampleCount);
    fb = fa;
  }
  return float(diffuseFresnel);
}

// Probe frame: shading normal +z, tangents from basis(+z), and the
// face-forward normal the closest-hit shaders pass to eval/pdf/sample.
static void localFrame(const V3<float>& V, V3<float>& N, V3<float>& X, V3<float>& Y) {
  N = V3<float>(0, 0, 1);
  basis(N, X, Y);
  if (!(dot(N, V) > 0)) N = -N;
}

void orc_bsdf_eval(const GpuMaterial* m, const float Lp[3], const float Vp[3], uint32_t flags, float out[3]) {
  V3<float> N, X, Y, L = load3<float>(Lp), V = load3<float>(Vp), r(0);
  localFrame(V, N, X, Y);
  V3<float> kd = load3<float>(m->diffuse);
  switch (m->type) {
    case MaterialTypeBrdfLambertian: r = lamEval(L, V, N, kd, flags); break;
    case MaterialTypeBrdfPbrMetalnessRoughness: {
      float ax = std::max(sqr(m->roughness), 0.001f);
      r = pbrEval(L, V, N, X, Y, kd, ax, ax, m->metalness, m->ior);
      break;
    }
    case MaterialTypeBrdfRoughPlastic: {
      float eta = m->ior, inv = 1.f / (eta * eta);
      float ax = std::max(0.001f, m->anisoAlpha[0]), ay = std::max(0.001f, m->anisoAlpha[1]);
      r = rpEval(L, V, N, X, Y, kd, V3<float>(1), eta, m->radiance[0], inv, ax, ay, flags);
      break;
    }
    case MaterialTypeBsdfDielectric: {
      float eta = V.z > 0 ? 1.f / m->ior : m->ior;
      r = dielEval(L, V, N, eta, flags);
      break;
    }
    case MaterialTypeBrdfMirror:
      if ((flags & EDelta) && std::fabs(dot(reflect(-V, N), L) - 1.f) < K<float>::EPS) r = kd;
      break;
    case MaterialTypeBrdfConductor: {
      float NdotL = dot(L, N), NdotV = dot(V, N);
      if (!(NdotL < 0 || NdotV < 0 || (flags & EDelta) == 0) && std::fabs(dot(reflect(-V, N), L) - 1.f) < K<float>::EPS)
        r = kd * conductorReflectance3(load3<float>(m->radiance), load3<float>(m->radianceFactor), NdotL);
      break;
    }
    case MaterialTypeBrdfRoughConductor: {
      float ax = std::max(m->anisoAlpha[0], K<float>::EPS), ay = std::max(m->anisoAlpha[1], K<float>::EPS);
      r = roughConductorEval(L, V, N, X, Y, kd, load3<float>(m->radiance), load3<float>(m->radianceFactor), ax, ay, flags);
      break;
    }
    case MaterialTypeBrdfPlastic: {
      float eta = m->ior;
      r = plasticEval(L, V, N, kd, eta, m->radiance[0], 1.f / (eta * eta), flags);
      break;
    }
    case MaterialTypeBrdfPhong: r = phongEval(L, V, N, kd, load3<float>(m->rhoSpec), m->specular, flags); break;
    case MaterialTypeBrdfKang18: {
      float ax = std::max(m->anisoAlpha[0], K<float>::EPS), ay = std::max(m->anisoAlpha[1], K<float>::EPS);
      r = kangEval(L, V, N, X, Y, kd, load3<float>(m->rhoSpec), ax, ay, m->ior);
      break;
    }
    case MaterialTypeBrdfDisney: {
      DisneyMat<float> dm = disneyMat<float>(*m, m->roughness);
      float eta = V.z > 0 ? 1.f / dm.ior : dm.ior, p;
      r = dsnEval(L, V, N, X, Y, dm, eta, flags, p);
      break;
    }
    default: break;
  }
  out[0] = r.x, out[1] = r.y, out[2] = r.z;
}

float orc_bsdf_pdf(const GpuMaterial* m, const float Lp[3], const float Vp[3], uint32_t flags) {
  V3<float> N, X, Y, L = load3<float>(Lp), V = load3<float>(Vp);
  localFrame(V, N, X, Y);
  switch (m->type) {
    case MaterialTypeBrdfLambertian: return lamPdf(L, N, flags);
    case MaterialTypeBrdfPbrMetalnessRoughness: {
      float ax = std::max(sqr(m->roughness), 0.001f);
      return pbrPdf(L, V, N, X, Y, ax, ax, flags);
    }
    case MaterialTypeBrdfRoughPlastic: {
      float ax = std::max(0.001f, m->anisoAlpha[0]), ay = std::max(0.001f, m->anisoAlpha[1]);
      return rpPdf(L, V, N, X, Y, ax, ay, m->ior, luminance(load3<float>(m->diffuse)), flags);
    }
    case MaterialTypeBsdfDielectric: {
      float eta = V.z > 0 ? 1.f / m->ior : m->ior;
      return dielPdf(L, V, N, eta, flags);
    }
    case MaterialTypeBrdfMirror:
    case MaterialTypeBrdfConductor:
      return (flags & EDelta) && dot(L, N) >= 0 && dot(V, N) >= 0 && std::fabs(dot(reflect(-V, N), L) - 1.f) < K<float>::EPS
                 ? 1.f
                 : 0.f;
    case MaterialTypeBrdfRoughConductor: {
      float ax = std::max(m->anisoAlpha[0], K<float>::EPS), ay = std::max(m->anisoAlpha[1], K<float>::EPS);
      if (dot(L, N) < 0 || dot(V, N) < 0 || (flags & EArea) == 0) return 0;
      V3<float> wi = toLocal(X, Y, N, L), wo = toLocal(X, Y, N, V);
      return importanceAnisoGGXPdf(makeNormal(wi + wo), wo, ax, ay);
    }
    case MaterialTypeBrdfPlastic: {
      V3<float> kd = load3<float>(m->diffuse);
      float dAvg = luminance(kd), sAvg = luminance(V3<float>(1));
      return plasticPdf(L, V, N, m->ior, sAvg / (dAvg + sAvg), flags);
    }
    case MaterialTypeBrdfPhong:
      return phongPdf(L, V, N, load3<float>(m->diffuse), load3<float>(m->rhoSpec), m->specular, flags);
    case MaterialTypeBrdfKang18: {
      float ax = std::max(m->anisoAlpha[0], K<float>::EPS), ay = std::max(m->anisoAlpha[1], K<float>::EPS);
      return kangPdf(L, V, N, X, Y, ax, ay, load3<float>(m->diffuse), load3<float>(m->rhoSpec), flags);
    }
    case MaterialTypeBrdfDisney: {
      DisneyMat<float> dm = disneyMat<float>(*m, m->roughness);
      float eta = V.z > 0 ? 1.f / dm.ior : dm.ior, p;
      dsnEval(L, V, N, X, Y, dm, eta, flags, p);
      return p;
    }
    default: return 0;
  }
}

float orc_conductor_reflectance(float eta, float k, float cosThetaI) { return conductorReflectance(eta, k, cosThetaI); }

// The rough-plastic light-sample terms exactly as chsRoughPlastic forms them
// (permuted eval and pdf), in the probe frame.
void orc_rough_plastic_light_terms(const GpuMaterial* m, const float Lp[3], const float Vp[3], uint32_t lightFlags,
                                   float bw[3], float* bsdfPdf) {
  V3<float> N, X, Y, L = load3<float>(Lp), V = load3<float>(Vp), w;
  localFrame(V, N, X, Y);
  const float eta = m->ior, invEta2 = 1.f / (eta * eta);
  const float ax = std::max(K<float>::EPS, m->anisoAlpha[0]), ay = std::max(K<float>::EPS, m->anisoAlpha[1]);
  const V3<float> kd = load3<float>(m->diffuse);
  rpLightTerms(L, V, N, X, Y, kd, eta, m->radiance[0], invEta2, ax, ay, luminance(kd), lightFlags, w, *bsdfPdf);
  bw[0] = w.x, bw[1] = w.y, bw[2] = w.z;
}

// sampleOneLight (sample_light.glsl:10-82) for one light at shading point p.
void orc_sample_light(const GpuLight* l, float r0, float r1, const float p[3], float rad[3], float d[3], float* dist,
                      float* pdf, uint32_t* flags) {
  LRec<float> rec{};
  V3<float> r = sampleOneLight<float>(r0, r1, *l, load3<float>(p), rec);
  rad[0] = r.x, rad[1] = r.y, rad[2] = r.z;
  d[0] = rec.d.x, d[1] = rec.d.y, d[2] = rec.d.z;
  *dist = rec.dist;
  *pdf = rec.pdf;
  *flags = rec.flags;
}

void orc_set_sun_and_sky(orc_scene* s, const GpuSunAndSky* sky) { s->sky = *sky; }
void orc_sun_and_sky(const GpuSunAndSky* sky, const float dir[3], int fp64, float out[3]) {
  if (fp64) {
    V3<double> r = sunAndSky(*sky, V3<double>(dir[0], dir[1], dir[2]));
    out[0] = float(r.x), out[1] = float(r.y), out[2] = float(r.z);
  } else {
    V3<float> r = sunAndSky(*sky, load3<float>(dir));
    out[0] = r.x, out[1] = r.y, out[2] = r.z;
  }
}
void orc_sky_physical_scale(float disk_scale, float glow_intensity, float disk_intensity, float out[2]) {
  float a, b;
  skyPhysicalScale(disk_scale, glow_intensity, disk_intensity, a, b);
  out[0] = a, out[1] = b;
}

void orc_bsdf_sample(const GpuMaterial* m, const float Vp[3], uint32_t* seed, float w[3], float dir[3], float* pdf,
                     uint32_t* flags) {
  V3<float> N, X, Y, V = load3<float>(Vp), d(0), r(0);
  localFrame(V, N, X, Y);
  V3<float> kd = load3<float>(m->diffuse);
  float p = 0;
  uint32_t f = 0;
  switch (m->type) {
    case MaterialTypeBrdfLambertian: {
      float u0 = rnd<float>(*seed), u1 = rnd<float>(*seed);
      V3<float> wi = cosineSampleHemisphere(u0, u1);
      p = cosineHemispherePdf(wi.z);
      d = toWorld(X, Y, N, wi);
      f = EDiffuseReflection;
      r = kd * K<float>::INV_PI * std::fabs(wi.z);
      break;
    }
    case MaterialTypeBrdfPbrMetalnessRoughness: {
      float ax = std::max(sqr(m->roughness), 0.001f);
      float u0 = rnd<float>(*seed), u1 = rnd<float>(*seed);
      r = pbrSample(u0, u1, *seed, V, N, X, Y, kd, ax, ax, m->metalness, m->ior, d, p, f);
      break;
    }
    case MaterialTypeBrdfRoughPlastic: {
      float eta = m->ior, inv = 1.f / (eta * eta);
      float ax = std::max(0.001f, m->anisoAlpha[0]), ay = std::max(0.001f, m->anisoAlpha[1]);
      float u0 = rnd<float>(*seed), u1 = rnd<float>(*seed);
      r = rpSample(u0, u1, V, N, X, Y, kd, V3<float>(1), eta, m->radiance[0], inv, ax, ay, luminance(kd), d, p, f);
      break;
    }
    case MaterialTypeBsdfDielectric: {
      float eta = V.z > 0 ? 1.f / m->ior : m->ior;
      float u = rnd<float>(*seed);
      r = dielSample(u, V, N, eta, d, p, f);
      break;
    }
    case MaterialTypeBrdfDisney: {
      DisneyMat<float> dm = disneyMat<float>(*m, m->roughness);
      float eta = V.z > 0 ? 1.f / dm.ior : dm.ior;
      float u0 = rnd<float>(*seed), u1 = rnd<float>(*seed);
      r = dsnSample(u0, u1, V, N, X, Y, dm, eta, d, p, f);
      break;
    }
    default: break;
  }
  w[0] = r.x, w[1] = r.y, w[2] = r.z;
  dir[0] = d.x, dir[1] = d.y, dir[2] = d.z;
  *pdf = p;
  *flags = f;
}

void orc_sample_envmap(const orc_scene* s, const float envT[16], float intensity, float r0, float r1, float L[3],
                       float* pdf, float rad[3]) {
  V3<float> l;
  float p;
  V3<float> r = sampleEnvmap<float>(*s, r0, r1, envT, float(s->env.w), float(s->env.h), intensity, l, p);
  L[0] = l.x, L[1] = l.y, L[2] = l.z;
  *pdf = p;
  rad[0] = r.x, rad[1] = r.y, rad[2] = r.z;
}
float orc_pdf_envmap(const orc_scene* s, const float envT[16], const float L[3]) {
  return pdfEnvmap<float>(*s, envT, float(s->env.w), float(s->env.h), load3<float>(L));
}
void orc_eval_envmap(const orc_scene* s, const float envT[16], float intensity, const float L[3], float rad[3]) {
  V3<float> r = evalEnvmap<float>(*s, envT, intensity, load3<float>(L));
  rad[0] = r.x, rad[1] = r.y, rad[2] = r.z;
}

}  // extern "C"
