// ORACLE / TEST INFRASTRUCTURE ONLY.
// refcpu: CPU restatement of RadiantiumPlusPlus' `path` / `vol_path` hot path,
// used by tests/ and by bench.py's cpu_baseline leg as the checker.  Nothing
// in the shipped library links or calls it.
//
// Float vector / matrix layer replacing Eigen in the reference.  Operations
// that the reference writes with Math::Fmadd use std::fma here as well
// (core/include/rad/core/math_base.h:68-72); everything else is plain float.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <utility>

namespace ref {

using Float = float;
constexpr Float PI = static_cast<Float>(3.1415926535897932);        // math_ext.h:8
constexpr Float Float32Epsilon = 0x1p-24f;                           // math_ext.h:9
constexpr Float RayEpsilon = Float32Epsilon * Float(1500);           // math_ext.h:12
constexpr Float ShadowEpsilon = RayEpsilon * Float(10);              // math_ext.h:16
constexpr Float OneMinusEpsilon = 0x1.fffffep-1f;                    // math_ext.h:18

inline Float Fmadd(Float a, Float b, Float c) { return std::fma(a, b, c); }
inline Float Sqr(Float v) { return v * v; }
inline Float Rcp(Float v) { return Float(1) / v; }
inline Float Rsqrt(Float v) { return Rcp(std::sqrt(v)); }
inline Float Sign(Float v) { return v >= 0 ? Float(1) : Float(-1); }
inline Float MulSign(Float a, Float b) { return b >= 0 ? a : -a; }
inline Float SafeSqrt(Float v) { return std::sqrt(std::max(v, Float(0))); }
inline Float Lerp(Float a, Float b, Float t) { return Fmadd(b, t, Fmadd(-a, t, a)); }
inline Float Radian(Float deg) { return deg * (PI / Float(180)); }

struct V2 {
  Float x = 0, y = 0;
  V2() = default;
  V2(Float a, Float b) : x(a), y(b) {}
  Float operator[](int i) const { return i == 0 ? x : y; }
};

struct V3 {
  Float x = 0, y = 0, z = 0;
  V3() = default;
  V3(Float a, Float b, Float c) : x(a), y(b), z(c) {}
  explicit V3(Float a) : x(a), y(a), z(a) {}
  Float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  Float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  V3 operator-() const { return {-x, -y, -z}; }
  V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
  V3& operator*=(Float s) { x *= s; y *= s; z *= s; return *this; }
  V3& operator/=(Float s) { x /= s; y /= s; z /= s; return *this; }
};

inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(const V3& a, Float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(Float s, const V3& a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(const V3& a, Float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 Mul(const V3& a, const V3& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 Div(const V3& a, const V3& b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline Float Dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 Cross(const V3& a, const V3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline Float SquaredNorm(const V3& a) { return Dot(a, a); }
inline Float Norm(const V3& a) { return std::sqrt(SquaredNorm(a)); }
inline V3 Normalized(const V3& a) {
  Float n = Norm(a);
  return n > 0 ? a / n : a;  // Eigen::normalized leaves a zero vector untouched
}
inline Float MaxCoeff(const V3& a) { return std::max(a.x, std::max(a.y, a.z)); }
inline Float MinCoeff(const V3& a) { return std::min(a.x, std::min(a.y, a.z)); }
inline V3 Abs(const V3& a) { return {std::abs(a.x), std::abs(a.y), std::abs(a.z)}; }
inline bool IsZero(const V3& a) { return a.x == 0 && a.y == 0 && a.z == 0; }
inline V3 Fmadd(const V3& a, const V3& b, const V3& c) {
  return {Fmadd(a.x, b.x, c.x), Fmadd(a.y, b.y, c.y), Fmadd(a.z, b.z, c.z)};
}
inline V3 MulSign(const V3& v, Float s) { return s >= 0 ? v : -v; }
inline Float AbsDot(const V3& a, const V3& b) { return std::abs(Dot(a, b)); }
inline V3 Min(const V3& a, const V3& b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V3 Max(const V3& a, const V3& b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline V3 Exp(const V3& a) { return {std::exp(a.x), std::exp(a.y), std::exp(a.z)}; }
inline V3 Sqrt(const V3& a) { return {std::sqrt(a.x), std::sqrt(a.y), std::sqrt(a.z)}; }

// RGB spectrum (spectrum.h:12-83): a float3.
using Spectrum = V3;
inline bool HasNaN(const Spectrum& s) { return std::isnan(s.x) || std::isnan(s.y) || std::isnan(s.z); }
inline Float Luminance(const Spectrum& s) {
  return (s.x * Float(0.212671)) + (s.y * Float(0.715160)) + (s.z * Float(0.072169));
}

// CoordinateSystem (math_ext.h:48-61).
inline std::pair<V3, V3> CoordinateSystem(const V3& n) {
  Float sign = Sign(n.z);
  Float a = -1 / (sign + n.z);
  Float b = n.x * n.y * a;
  V3 X(MulSign(Sqr(n.x) * a, n.z) + 1.0f, MulSign(b, n.z), MulSign(n.x, -n.z));
  V3 Y(b, sign + Sqr(n.y) * a, -n.y);
  return {X, Y};
}

struct Mat4 {
  Float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
  static Mat4 FromRowMajor(const float* a) {
    Mat4 r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) r.m[i][j] = a[i * 4 + j];
    return r;
  }
  Mat4 operator*(const Mat4& o) const {
    Mat4 r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        Float s = 0;
        for (int k = 0; k < 4; k++) s += m[i][k] * o.m[k][j];
        r.m[i][j] = s;
      }
    return r;
  }
  bool IsIdentity() const {
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++)
        if (std::abs(m[i][j] - (i == j ? 1.f : 0.f)) > 1e-12f) return false;
    return true;
  }
  // Cofactor inverse (computeInverseWithCheck in transform.cpp:9-16).
  bool Inverse(Mat4& out) const {
    const Float* a = &m[0][0];
    Float inv[16];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    Float det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    if (det == 0 || !std::isfinite(det)) return false;
    Float invDet = Float(1) / det;
    for (int i = 0; i < 16; i++) (&out.m[0][0])[i] = inv[i] * invDet;
    return true;
  }
};

// Transform (transform.cpp:5-95).
struct Transform {
  Mat4 ToWorld, ToLocal;
  Transform() = default;
  explicit Transform(const Mat4& toWorld) : ToWorld(toWorld) {
    if (!toWorld.Inverse(ToLocal)) throw std::runtime_error("matrix not invertible");
  }
  static V3 Affine(const Mat4& M, const V3& v) {
    Float r[4];
    for (int i = 0; i < 4; i++) r[i] = M.m[i][0] * v.x + M.m[i][1] * v.y + M.m[i][2] * v.z + M.m[i][3];
    return V3(r[0], r[1], r[2]) / r[3];
  }
  static V3 Linear(const Mat4& M, const V3& v) {
    return {M.m[0][0] * v.x + M.m[0][1] * v.y + M.m[0][2] * v.z,
            M.m[1][0] * v.x + M.m[1][1] * v.y + M.m[1][2] * v.z,
            M.m[2][0] * v.x + M.m[2][1] * v.y + M.m[2][2] * v.z};
  }
  V3 ApplyAffineToWorld(const V3& v) const { return Affine(ToWorld, v); }
  V3 ApplyLinearToWorld(const V3& v) const { return Linear(ToWorld, v); }
  // linear(ToLocal)^T * v (transform.cpp:68-70)
  V3 ApplyNormalToWorld(const V3& v) const {
    const auto& L = ToLocal.m;
    return {L[0][0] * v.x + L[1][0] * v.y + L[2][0] * v.z,
            L[0][1] * v.x + L[1][1] * v.y + L[2][1] * v.z,
            L[0][2] * v.x + L[1][2] * v.y + L[2][2] * v.z};
  }
  V3 ApplyAffineToLocal(const V3& v) const { return Affine(ToLocal, v); }
  V3 ApplyLinearToLocal(const V3& v) const { return Linear(ToLocal, v); }
  V3 TranslationToWorld() const { return {ToWorld.m[0][3], ToWorld.m[1][3], ToWorld.m[2][3]}; }
  bool IsIdentity() const { return ToWorld.IsIdentity(); }
  bool HasNonUniformScale() const {
    Float la2 = SquaredNorm(ApplyLinearToWorld(V3(1, 0, 0)));
    Float lb2 = SquaredNorm(ApplyLinearToWorld(V3(0, 1, 0)));
    Float lc2 = SquaredNorm(ApplyLinearToWorld(V3(0, 0, 1)));
    return std::abs(la2 - lb2) > 0.0001f || std::abs(la2 - lc2) > 0.0001f;
  }
};

// Frame (frame.cpp:9-72).
struct Frame {
  V3 S, T, N;
  Frame() = default;
  explicit Frame(const V3& n) {
    auto st = CoordinateSystem(n);
    S = st.first;
    T = st.second;
    N = n;
  }
  Frame(const V3& s, const V3& t, const V3& n) : S(s), T(t), N(n) {}
  V3 ToWorld(const V3& v) const { return N * v.z + (T * v.y + (S * v.x)); }
  V3 ToLocal(const V3& v) const { return {Dot(v, S), Dot(v, T), Dot(v, N)}; }
  static Float CosTheta(const V3& v) { return v.z; }
  static Float AbsCosTheta(const V3& v) { return std::abs(v.z); }
  static Float Sin2Theta(const V3& v) { return Fmadd(v.x, v.x, Sqr(v.y)); }
  static bool IsSameHemisphere(const V3& a, const V3& b) { return CosTheta(a) * CosTheta(b) > 0; }
  // frame.cpp:48-56.  With quirks on, the clamp `cwiseMin(-1).cwiseMax(1)`
  // maps every off-normal input to (sin, cos) = (1, 1); with quirks off the
  // intended clamp to [-1, 1] is used.
  static std::pair<Float, Float> SinCosPhi(const V3& v, bool quirks) {
    Float sin2Theta = Sin2Theta(v);
    Float invSinTheta = Rsqrt(Sin2Theta(v));
    Float rx = v.x * invSinTheta, ry = v.y * invSinTheta;
    if (std::abs(sin2Theta) <= 4 * Float32Epsilon) {
      rx = 1;
      ry = 0;
    } else if (quirks) {
      rx = std::max(std::min(rx, Float(-1)), Float(1));
      ry = std::max(std::min(ry, Float(-1)), Float(1));
    } else {
      rx = std::min(std::max(rx, Float(-1)), Float(1));
      ry = std::min(std::max(ry, Float(-1)), Float(1));
    }
    return {ry, rx};
  }
};

struct Ray {
  V3 O, D;
  Float MinT = 0, MaxT = std::numeric_limits<Float>::max();
  V3 operator()(Float t) const { return D * t + O; }
};

}  // namespace ref
