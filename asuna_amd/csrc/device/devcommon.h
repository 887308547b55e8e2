// Device-side building blocks of the render path for gfx950: vector math,
// RNG, warps, light and BxDF evaluation.  Every function states the GLSL
// function it stands in for; evaluation order follows the GLSL expression
// so that FP32 results round the same way (the library is compiled with
// -ffp-contract=off; fused multiply-adds appear only where written).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/asuna_hip.h"

namespace ashd {

// ------------------------------------------------------------ constants
// math.glsl:4-15 (float literals)
constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv2Pi = 0.15915494309189533577f;
constexpr float kInv4Pi = 0.07957747154594766788f;
constexpr float kPiOver2 = 1.57079632679489661923f;
constexpr float kPiOver4 = 0.78539816339744830961f;
constexpr float kEps = 0.001f;
constexpr float kInf = 10000000000.0f;
constexpr float kMinimum = 0.00001f;

// structs.glsl:16-46
constexpr uint32_t kBsdfNull = 0, kDiffuseRefl = 1, kGlossyRefl = 4, kSpecRefl = 16, kSpecTrans = 32;
constexpr uint32_t kSmooth = 1 | 2 | 4 | 8;
constexpr uint32_t kLightDelta = 1, kLightArea = 2;
__device__ __forceinline__ bool isNonSpecular(uint32_t f) { return (f & kSmooth) != 0; }

// ------------------------------------------------------------------ vec3
struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 mk1(float a) { return f3{a, a, a}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
// normalize(v) = v * (1 / |v|) (one correctly rounded sqrt and divide).
__device__ __forceinline__ f3 normalize(f3 a) {
  float inv = 1.0f / sqrtf(dot(a, a));
  return a * inv;
}
// makeNormal, math.glsl:58-61
__device__ __forceinline__ f3 makeNormal(f3 n) {
  if (length(n) == 0.0f) return n;
  return normalize(n);
}
__device__ __forceinline__ f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float sqr(float x) { return x * x; }

// Column-major 4x4 (GLSL mat4): transformPoint / transformVector /
// transformDirection, math.glsl:46-66.
__device__ __forceinline__ f3 transformPoint(const float* m, f3 p) {
  float x = m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12];
  float y = m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13];
  float z = m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14];
  float w = m[3] * p.x + m[7] * p.y + m[11] * p.z + m[15];
  return f3{x, y, z} / w;
}
__device__ __forceinline__ f3 transformVector(const float* m, f3 v) {
  return f3{m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
            m[2] * v.x + m[6] * v.y + m[10] * v.z};
}
__device__ __forceinline__ f3 transformDirection(const float* m, f3 d) { return makeNormal(transformVector(m, d)); }
// transpose(m) applied to a direction (pdfEnvmap / evalEnvmap).
__device__ __forceinline__ f3 transformVectorT(const float* m, f3 v) {
  return f3{m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z};
}

// ------------------------------------------------------------------- RNG
// xxhash32Seed (math.glsl:20-30), pcg (34-39), rand (42).
__device__ __forceinline__ uint32_t xxhash32Seed(uint32_t px, uint32_t py, uint32_t pz) {
  const uint32_t p0 = 2246822519U, p1 = 3266489917U, p2 = 668265263U, p3 = 374761393U;
  uint32_t h = pz + p3 + px * p1;
  h = p2 * ((h << 17) | (h >> 15));
  h += py * p1;
  h = p2 * ((h << 17) | (h >> 15));
  h = p0 * (h ^ (h >> 15));
  h = p1 * (h ^ (h >> 13));
  return h ^ (h >> 16);
}
__device__ __forceinline__ uint32_t pcg(uint32_t& s) {
  uint32_t prev = s * 747796405u + 2891336453u;
  uint32_t word = ((prev >> ((prev >> 28u) + 4u)) ^ prev) * 277803737u;
  s = prev;
  return (word >> 22u) ^ word;
}
// float(u) * (1/float(0xffffffff)) = float(u) * 2^-32
__device__ __forceinline__ float rnd(uint32_t& s) { return float(pcg(s)) * 2.3283064365386963e-10f; }

// ------------------------------------------------------------ warps etc.
__device__ __forceinline__ f3 toWorld(f3 X, f3 Y, f3 Z, f3 v) { return v.x * X + v.y * Y + v.z * Z; }
__device__ __forceinline__ f3 toLocal(f3 X, f3 Y, f3 Z, f3 v) { return f3{dot(v, X), dot(v, Y), dot(v, Z)}; }
__device__ __forceinline__ f3 reflect(f3 i, f3 n) { return i - 2.0f * dot(n, i) * n; }
__device__ __forceinline__ f3 refract(f3 i, f3 n, float eta) {
  float d = dot(n, i);
  float k = 1.0f - eta * eta * (1.0f - d * d);
  if (k < 0.0f) return mk1(0.f);
  return eta * i - (eta * d + sqrtf(k)) * n;
}
// concentricSampleDisk, math.glsl:148-163
__device__ __forceinline__ void concentricDisk(float u0, float u1, float& ox, float& oy) {
  float ux = 2.0f * u0 - 1.0f, uy = 2.0f * u1 - 1.0f;
  if (ux == 0.0f && uy == 0.0f) {
    ox = oy = 0.f;
    return;
  }
  float theta, r;
  if (fabsf(ux) > fabsf(uy)) {
    r = ux;
    theta = kPiOver4 * (uy / ux);
  } else {
    r = uy;
    theta = kPiOver2 - kPiOver4 * (ux / uy);
  }
  ox = r * cosf(theta);
  oy = r * sinf(theta);
}
// cosineSampleHemisphere / cosineHemispherePdf, math.glsl:165-174
__device__ __forceinline__ f3 cosineHemisphere(float u0, float u1) {
  float dx, dy;
  concentricDisk(u0, u1, dx, dy);
  float z = sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy));
  return f3{dx, dy, z};
}
__device__ __forceinline__ float cosinePdf(float c) { return c <= 0.0f ? 0.0f : c * kInvPi; }
// uniformSampleSphere, math.glsl:133-138 (pdf 1 / (4 pi))
__device__ __forceinline__ f3 uniformSphere(float u0, float u1) {
  float z = 1.0f - 2.0f * u0;
  float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
  float phi = kTwoPi * u1;
  return f3{r * cosf(phi), r * sinf(phi), z};
}
// powerHeuristic, math.glsl:184-189
__device__ __forceinline__ float powerHeuristic(float a, float b) {
  a = a * a;
  b = b * b + a;
  if (b == 0.0f) return 0.0f;
  return a / b;
}
// basis, math.glsl:199-216
__device__ __forceinline__ void basis(f3 n, f3& f, f3& r) {
  if (n.z < -0.999999f) {
    f = f3{0.f, -1.f, 0.f};
    r = f3{-1.f, 0.f, 0.f};
  } else {
    float a = 1.0f / (1.0f + n.z);
    float b = -n.x * n.y * a;
    f = f3{1.0f - n.x * n.x * a, b, -n.x};
    r = f3{b, 1.0f - n.y * n.y * a, -n.y};
  }
}
// offsetPositionAlongNormal, math.glsl:241-266
__device__ __forceinline__ float offComp(float p, float n) {
  int of = int(256.0f * n);
  int bits = __float_as_int(p) + ((p < 0.0f) ? -of : of);
  return fabsf(p) < (1.0f / 32.0f) ? p + (1.0f / 65536.0f) * n : __int_as_float(bits);
}
__device__ __forceinline__ f3 offsetRay(f3 p, f3 n) { return f3{offComp(p.x, n.x), offComp(p.y, n.y), offComp(p.z, n.z)}; }
// luminance, sun_and_sky.glsl:29-31
__device__ __forceinline__ float luminance(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// --------------------------------------------------- texture emulation
// texture(sampler2D, uv) at LOD 0, linear filter, REPEAT addressing
// (texture.cpp:99-130, 240-296; Vulkan "Texel Filtering").
struct DevImage {
  const float4* px;
  uint32_t w, h;
};
__device__ __forceinline__ int wrapi(int i, int n) {
  int m = i % n;
  return m < 0 ? m + n : m;
}
__device__ __forceinline__ float4 tex2D(const DevImage& im, float u, float v) {
  if (!(u == u)) u = 0.f;
  if (!(v == v)) v = 0.f;
  float x = u * float(im.w) - 0.5f, y = v * float(im.h) - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float a = x - fx, b = y - fy;
  int ix = int(fx), iy = int(fy);
  int x0 = wrapi(ix, int(im.w)), y0 = wrapi(iy, int(im.h));
  int x1 = wrapi(ix + 1, int(im.w)), y1 = wrapi(iy + 1, int(im.h));
  float4 t00 = im.px[size_t(y0) * im.w + x0], t10 = im.px[size_t(y0) * im.w + x1];
  float4 t01 = im.px[size_t(y1) * im.w + x0], t11 = im.px[size_t(y1) * im.w + x1];
  float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
  return make_float4(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
                     w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
                     w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z,
                     w00 * t00.w + w10 * t10.w + w01 * t01.w + w11 * t11.w);
}
// The marginal table holds data in column 0 only (texture.cpp:209-214); a
// lookup at u = 0 blends column 0 with column W-1 (zero) through REPEAT.
__device__ __forceinline__ float2 texMarginal(const float2* m, uint32_t W, uint32_t H, float v) {
  if (!(v == v)) v = 0.f;
  float x = 0.0f * float(W) - 0.5f, y = v * float(H) - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float a = x - fx, b = y - fy;
  int ix = int(fx), iy = int(fy);
  int x0 = wrapi(ix, int(W)), x1 = wrapi(ix + 1, int(W));
  int y0 = wrapi(iy, int(H)), y1 = wrapi(iy + 1, int(H));
  float2 z = make_float2(0.f, 0.f);
  float2 t00 = x0 == 0 ? m[y0] : z, t10 = x1 == 0 ? m[y0] : z;
  float2 t01 = x0 == 0 ? m[y1] : z, t11 = x1 == 0 ? m[y1] : z;
  float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
  return make_float2(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
                     w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y);
}
// The env map's importance tables.  The conditional table's two channels
// are stored apart: its inverse-CDF channel as a float plane, its pdf
// channel in the env texels' alpha, so that the pdf and the radiance of one
// direction come from the same four texel loads (the two textures have the
// same size, hence the same footprint and weights).  Both planes are laid
// out in tiles of one 128-byte line (env: 4x2 texels of 16 B; inverse CDF:
// 8x4 texels of 4 B), so that a bilinear footprint touches about 1.9 and
// 1.4 lines instead of 2.25 and 2.06 in row-major order.  Each channel is
// filtered exactly as a two-channel row-major lookup filters it.
__host__ __device__ inline uint32_t envTilesX(uint32_t W) { return (W + 3) / 4; }
__host__ __device__ inline uint32_t condTilesX(uint32_t W) { return (W + 7) / 8; }
__host__ __device__ inline size_t envTexel(uint32_t W, uint32_t x, uint32_t y) {
  return ((size_t(y >> 1) * envTilesX(W) + (x >> 2)) << 3) | ((y & 1u) << 2) | (x & 3u);
}
__host__ __device__ inline size_t condTexel(uint32_t W, uint32_t x, uint32_t y) {
  return ((size_t(y >> 2) * condTilesX(W) + (x >> 3)) << 5) | ((y & 3u) << 3) | (x & 7u);
}
__host__ __device__ inline size_t envTexels(uint32_t W, uint32_t H) { return size_t(envTilesX(W)) * 8 * ((H + 1) / 2); }
__host__ __device__ inline size_t condTexels(uint32_t W, uint32_t H) { return size_t(condTilesX(W)) * 32 * ((H + 3) / 4); }

__device__ __forceinline__ float texCondX(const float* c, uint32_t W, uint32_t H, float u, float v) {
  if (!(u == u)) u = 0.f;
  if (!(v == v)) v = 0.f;
  float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float a = x - fx, b = y - fy;
  int ix = int(fx), iy = int(fy);
  uint32_t x0 = wrapi(ix, int(W)), y0 = wrapi(iy, int(H));
  uint32_t x1 = wrapi(ix + 1, int(W)), y1 = wrapi(iy + 1, int(H));
  float t00 = c[condTexel(W, x0, y0)], t10 = c[condTexel(W, x1, y0)];
  float t01 = c[condTexel(W, x0, y1)], t11 = c[condTexel(W, x1, y1)];
  float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
  return w00 * t00 + w10 * t10 + w01 * t01 + w11 * t11;
}
// tex2D on the tiled env plane
__device__ __forceinline__ float4 texEnv(const DevImage& im, float u, float v) {
  if (!(u == u)) u = 0.f;
  if (!(v == v)) v = 0.f;
  float x = u * float(im.w) - 0.5f, y = v * float(im.h) - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float a = x - fx, b = y - fy;
  int ix = int(fx), iy = int(fy);
  uint32_t x0 = wrapi(ix, int(im.w)), y0 = wrapi(iy, int(im.h));
  uint32_t x1 = wrapi(ix + 1, int(im.w)), y1 = wrapi(iy + 1, int(im.h));
  float4 t00 = im.px[envTexel(im.w, x0, y0)], t10 = im.px[envTexel(im.w, x1, y0)];
  float4 t01 = im.px[envTexel(im.w, x0, y1)], t11 = im.px[envTexel(im.w, x1, y1)];
  float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
  return make_float4(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
                     w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
                     w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z,
                     w00 * t00.w + w10 * t10.w + w01 * t01.w + w11 * t11.w);
}

// ---------------------------------------------------------- microfacet
// dAnisoGGX / importanceSampleAnisoGGX / importanceAnisoGGXPdf /
// g1SmithAnisoGGX (brdf_pbr_metalness_roughness.rchit:18-53, identical in
// brdf_rough_plastic.rchit:45-80).
__device__ __forceinline__ float dAnisoGGX(float HdotN, float HdotX, float HdotY, float ax, float ay) {
  return 1.0f / (kPi * ax * ay * sqr(sqr(HdotX / ax) + sqr(HdotY / ay) + sqr(HdotN)) + kEps);
}
__device__ __forceinline__ f3 sampleGGX(float u1, float u2, f3 wo, float ax, float ay) {
  float factor = sqrtf(fmaxf(0.0f, u1 / fmaxf(1.0f - u1, kEps)));
  float phi = kTwoPi * u2;
  f3 wh = f3{0.f, 0.f, 1.f};
  wh.x = -ax * factor * cosf(phi);
  wh.y = -ay * factor * sinf(phi);
  wh = makeNormal(wh);
  return reflect(-wo, wh);
}
__device__ __forceinline__ float pdfGGX(f3 wh, f3 wo, float ax, float ay) {
  float pdf = 0.0f, HdotV = dot(wh, wo);
  f3 wi = reflect(-wo, wh);
  if (wi.z > 0.0f && wo.z > 0.0f && wh.z > 0.0f)
    pdf = dAnisoGGX(wh.z, wh.x, wh.y, ax, ay) * fabsf(wh.z) / (4.0f * HdotV + kEps);
  return pdf;
}
__device__ __forceinline__ float g1GGX(float NdotV, float VdotX, float VdotY, float ax, float ay) {
  if (NdotV <= 0.0f) return 0.0f;
  f3 f = f3{ax * VdotX, ay * VdotY, NdotV};
  return 1.0f / (NdotV + length(f));
}
// fresnelDielectricExt, brdf_rough_plastic.rchit:12-41
__device__ __forceinline__ float fresnelDielectricExt(float cosI_, float eta) {
  if (eta == 1.0f) return 0.0f;
  float scale = (cosI_ > 0.0f) ? 1.0f / eta : eta;
  float cosTSqr = 1.0f - (1.0f - cosI_ * cosI_) * (scale * scale);
  if (cosTSqr <= 0.0f) return 1.0f;
  float cosI = fabsf(cosI_);
  float cosT = sqrtf(cosTSqr);
  float Rs = (cosI - eta * cosT) / (cosI + eta * cosT);
  float Rp = (eta * cosI - cosT) / (eta * cosI + cosT);
  return 0.5f * (Rs * Rs + Rp * Rp);
}
// dielectricFresnel, bsdf_dielectric.rchit:12-24
__device__ __forceinline__ float dielectricFresnel(float cosI, float eta) {
  float sinTSq = eta * eta * (1.0f - cosI * cosI);
  if (sinTSq > 1.0f) return 1.0f;
  float cosT = sqrtf(fmaxf(1.0f - sinTSq, 0.0f));
  float rs = (eta * cosT - cosI) / (eta * cosT + cosI);
  float rp = (eta * cosI - cosT) / (eta * cosI + cosT);
  return 0.5f * (rs * rs + rp * rp);
}

}  // namespace ashd
