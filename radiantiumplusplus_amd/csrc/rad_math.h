// Float3 arithmetic usable from host and device code, with the reference's
// numerical conventions (Math::Fmadd = fused multiply-add, Rcp = 1/x, ...,
// module/rad.core/include/rad/core/math_base.h:37-104 and
// module/rad.offline/include/rad/offline/math_ext.h:8-61).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define RAD_HD __host__ __device__ __forceinline__

namespace rad {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kFloat32Eps = 0x1p-24f;
constexpr float kRayEps = kFloat32Eps * 1500.0f;     // math_ext.h:12
constexpr float kShadowEps = kRayEps * 10.0f;        // math_ext.h:16
constexpr float kOneMinusEps = 0x1.fffffep-1f;
constexpr float kInf = __builtin_huge_valf();
constexpr float kFloatMax = 3.402823466e+38f;

struct f3 {
  float x, y, z;
};

RAD_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
RAD_HD f3 splat(float v) { return f3{v, v, v}; }
RAD_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RAD_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RAD_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
RAD_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RAD_HD f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
RAD_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
RAD_HD f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
RAD_HD f3 div(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
RAD_HD f3& operator+=(f3& a, f3 b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  return a;
}
RAD_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RAD_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
RAD_HD float sqnorm(f3 a) { return dot(a, a); }
RAD_HD float norm(f3 a) { return sqrtf(sqnorm(a)); }
RAD_HD f3 normalize(f3 a) {
  float n = norm(a);
  return n > 0.f ? a / n : a;
}
RAD_HD float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
RAD_HD float minc(f3 a) { return fminf(a.x, fminf(a.y, a.z)); }
RAD_HD f3 absv(f3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
RAD_HD bool is_zero(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
RAD_HD float sqr(float v) { return v * v; }
RAD_HD float rcp(float v) { return 1.0f / v; }
RAD_HD float rsqrt_(float v) { return 1.0f / sqrtf(v); }
RAD_HD float signf_(float v) { return v >= 0.f ? 1.f : -1.f; }
RAD_HD float mulsign(float a, float b) { return b >= 0.f ? a : -a; }
RAD_HD f3 mulsign(f3 a, float b) { return b >= 0.f ? a : -a; }
RAD_HD float safe_sqrt(float v) { return sqrtf(fmaxf(v, 0.f)); }
RAD_HD float fmadd(float a, float b, float c) { return fmaf(a, b, c); }
RAD_HD f3 fmadd(f3 a, f3 b, f3 c) { return {fmaf(a.x, b.x, c.x), fmaf(a.y, b.y, c.y), fmaf(a.z, b.z, c.z)}; }
RAD_HD float lerp_(float a, float b, float t) { return fmaf(b, t, fmaf(-a, t, a)); }
RAD_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
RAD_HD f3 expv(f3 a) { return {expf(a.x), expf(a.y), expf(a.z)}; }
RAD_HD f3 sqrtv(f3 a) { return {sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)}; }
RAD_HD f3 minv(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
RAD_HD f3 maxv(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
RAD_HD float luminance(f3 s) { return (s.x * 0.212671f) + (s.y * 0.715160f) + (s.z * 0.072169f); }

// CoordinateSystem (math_ext.h:48-61)
RAD_HD void coordinate_system(f3 n, f3& X, f3& Y) {
  float sign = signf_(n.z);
  float a = -1.0f / (sign + n.z);
  float b = n.x * n.y * a;
  X = mk(mulsign(sqr(n.x) * a, n.z) + 1.0f, mulsign(b, n.z), mulsign(n.x, -n.z));
  Y = mk(b, sign + sqr(n.y) * a, -n.y);
}

// Orthonormal shading frame (frame.cpp:9-23).
struct Frame {
  f3 s, t, n;
  RAD_HD f3 to_world(f3 v) const { return n * v.z + (t * v.y + (s * v.x)); }
  RAD_HD f3 to_local(f3 v) const { return mk(dot(v, s), dot(v, t), dot(v, n)); }
};
RAD_HD Frame frame_from_n(f3 n) {
  Frame f;
  coordinate_system(n, f.s, f.t);
  f.n = n;
  return f;
}

// 3x4 affine matrix (row-major): rows r0, r1, r2 of a 4x4 whose last row is 0001.
struct Affine {
  float m[12];
  RAD_HD f3 point(f3 v) const {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3], m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
              m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]);
  }
  RAD_HD f3 vec(f3 v) const {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z);
  }
  // transpose of the linear part times v (normal transform from the inverse)
  RAD_HD f3 tvec(f3 v) const {
    return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
              m[2] * v.x + m[6] * v.y + m[10] * v.z);
  }
};

// Full 4x4 (camera clip transform needs the projective row).
struct Mat4 {
  float m[16];
  RAD_HD f3 point_h(f3 v) const {  // ApplyAffine* with the homogeneous divide (transform.cpp:39-42)
    float r0 = m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3];
    float r1 = m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7];
    float r2 = m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11];
    float r3 = m[12] * v.x + m[13] * v.y + m[14] * v.z + m[15];
    return mk(r0 / r3, r1 / r3, r2 / r3);
  }
};

}  // namespace rad
