// Host-side 4x4 float matrix helpers with nvmath semantics (column-major
// storage, element (r,c) at m[c*4+r]) for camera and instance transforms.
#pragma once

#include <cmath>
#include <cstring>

namespace ash {

struct Vec3f {
  float x = 0, y = 0, z = 0;
  Vec3f() = default;
  Vec3f(float a, float b, float c) : x(a), y(b), z(c) {}
};
inline Vec3f operator+(Vec3f a, Vec3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3f operator-(Vec3f a, Vec3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3f operator*(Vec3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot3(Vec3f a, Vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3f cross3(Vec3f a, Vec3f b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float len3(Vec3f a) { return std::sqrt(dot3(a, a)); }
// nvmath normalize: scale by 1/|v| (zero for a null vector).
inline Vec3f norm3(Vec3f a) {
  float n = len3(a);
  return n > 0.f ? a * (1.f / n) : Vec3f();
}

struct Mat4 {
  float m[16];
  Mat4() { identity(); }
  void identity() {
    std::memset(m, 0, sizeof m);
    m[0] = m[5] = m[10] = m[15] = 1.f;
  }
  static Mat4 zero() {
    Mat4 r;
    std::memset(r.m, 0, sizeof r.m);
    return r;
  }
  float& at(int r, int c) { return m[c * 4 + r]; }
  float at(int r, int c) const { return m[c * 4 + r]; }
};

inline Mat4 mul(const Mat4& a, const Mat4& b) {
  Mat4 o = Mat4::zero();
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      o.at(r, c) = a.at(r, 0) * b.at(0, c) + a.at(r, 1) * b.at(1, c) + a.at(r, 2) * b.at(2, c) +
                   a.at(r, 3) * b.at(3, c);
  return o;
}
inline Mat4 scaleMat(float x, float y, float z) {
  Mat4 o;
  o.at(0, 0) = x, o.at(1, 1) = y, o.at(2, 2) = z;
  return o;
}
inline Mat4 translateMat(float x, float y, float z) {
  Mat4 o;
  o.at(0, 3) = x, o.at(1, 3) = y, o.at(2, 3) = z;
  return o;
}
// Right-handed rotations about the principal axes (radians).
inline Mat4 rotX(float a) {
  Mat4 o;
  float c = std::cos(a), s = std::sin(a);
  o.at(1, 1) = c, o.at(1, 2) = -s, o.at(2, 1) = s, o.at(2, 2) = c;
  return o;
}
inline Mat4 rotY(float a) {
  Mat4 o;
  float c = std::cos(a), s = std::sin(a);
  o.at(0, 0) = c, o.at(0, 2) = s, o.at(2, 0) = -s, o.at(2, 2) = c;
  return o;
}
inline Mat4 rotZ(float a) {
  Mat4 o;
  float c = std::cos(a), s = std::sin(a);
  o.at(0, 0) = c, o.at(0, 1) = -s, o.at(1, 0) = s, o.at(1, 1) = c;
  return o;
}

// gluLookAt-style view matrix (world -> eye, camera looking down -z).
inline Mat4 lookAt(Vec3f eye, Vec3f center, Vec3f up) {
  Vec3f z = norm3(eye - center);
  Vec3f x = cross3(up, z);
  Vec3f y = cross3(z, x);
  x = norm3(x);
  y = norm3(y);
  Mat4 o;
  o.at(0, 0) = x.x, o.at(0, 1) = x.y, o.at(0, 2) = x.z;
  o.at(1, 0) = y.x, o.at(1, 1) = y.y, o.at(1, 2) = y.z;
  o.at(2, 0) = z.x, o.at(2, 1) = z.y, o.at(2, 2) = z.z;
  o.at(0, 3) = -x.x * eye.x - x.y * eye.y - x.z * eye.z;
  o.at(1, 3) = -y.x * eye.x - y.y * eye.y - y.z * eye.z;
  o.at(2, 3) = -z.x * eye.x - z.y * eye.y - z.z * eye.z;
  return o;
}

// Inverse of a rotation+translation matrix: [R^T | -R^T t].
inline Mat4 invertRotTrans(const Mat4& a) {
  Mat4 o;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o.at(r, c) = a.at(c, r);
  for (int r = 0; r < 3; ++r)
    o.at(r, 3) = -(a.at(0, r) * a.at(0, 3) + a.at(1, r) * a.at(1, 3) + a.at(2, r) * a.at(2, 3));
  return o;
}

// General inverse by cofactor expansion in float.
inline bool invertGeneral(const Mat4& in, Mat4& out) {
  const float* m = in.m;
  float inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  float det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  if (det == 0.f) return false;
  float idet = 1.0f / det;
  for (int i = 0; i < 16; ++i) out.m[i] = inv[i] * idet;
  return true;
}

}  // namespace ash
