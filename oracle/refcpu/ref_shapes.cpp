// ORACLE / TEST INFRASTRUCTURE ONLY (see ref_math.h).
// Shapes, lights, camera, scene assembly and the CPU accel of the restatement.
#include <immintrin.h>

#include <chrono>
#include <functional>
#include <future>

#include "ref_scene.h"

namespace ref {

// ---------------------------------------------------------------- triangles
// Products of the primitive tests with every rounding spelled out (fused
// multiply-adds in a fixed association): the device kernels evaluate the same
// expressions, so hits agree bit for bit.
static inline Float DotX(const V3& a, const V3& b) { return std::fma(a.x, b.x, std::fma(a.y, b.y, a.z * b.z)); }
static inline V3 CrossX(const V3& a, const V3& b) {
  return {std::fma(a.y, b.z, -(a.z * b.y)), std::fma(a.z, b.x, -(a.x * b.z)), std::fma(a.x, b.y, -(a.y * b.x))};
}

bool IntersectTriangleMT(const V3& v0, const V3& v1, const V3& v2, const Ray& ray, Float tnear, Float tfar,
                         Float& t, Float& u, Float& v) {
  V3 e1 = v0 - v1;
  V3 e2 = v2 - v0;
  V3 Ng = CrossX(e2, e1);
  V3 C = v0 - ray.O;
  V3 R = CrossX(C, ray.D);
  Float den = DotX(Ng, ray.D);
  Float absDen = std::abs(den);
  Float U = DotX(R, e2);
  Float V = DotX(R, e1);
  if (den < 0) { U = -U; V = -V; }
  if (!(den != 0 && U >= 0 && V >= 0 && U + V <= absDen)) return false;
  Float T = DotX(Ng, C);
  if (den < 0) T = -T;
  if (!(absDen * tnear < T && T <= absDen * tfar)) return false;
  t = T / absDen;
  u = U / absDen;
  v = V / absDen;
  return true;
}

Bounds MeshShape::PrimBounds(uint32_t prim) const {
  Bounds b;
  for (int k = 0; k < 3; k++) b.Extend(position[indices[prim * 3 + k]]);
  return b;
}

bool MeshShape::IntersectPrim(uint32_t prim, const Ray& ray, Float near, Float far, HitShapeRecord& rec) const {
  const V3& p0 = position[indices[prim * 3 + 0]];
  const V3& p1 = position[indices[prim * 3 + 1]];
  const V3& p2 = position[indices[prim * 3 + 2]];
  Float t, u, v;
  if (!IntersectTriangleMT(p0, p1, p2, ray, near, far, t, u, v)) return false;
  rec.T = t;
  rec.PrimitiveUV = V2(u, v);
  rec.PrimitiveIndex = prim;
  rec.ShapeIndex = index;
  rec.GeometryNormal = Cross(p1 - p0, p2 - p0);
  return true;
}

// mesh_base.cpp:38-82
SurfaceInteraction MeshShape::ComputeInteraction(const Ray& ray, const HitShapeRecord& rec) const {
  uint32_t face = rec.PrimitiveIndex * 3;
  uint32_t f0 = indices[face + 0], f1 = indices[face + 1], f2 = indices[face + 2];
  V3 p0 = position[f0], p1 = position[f1], p2 = position[f2];
  V2 primUV = rec.PrimitiveUV;
  V3 bary(1.f - primUV.x - primUV.y, primUV.x, primUV.y);
  V3 dp0 = p1 - p0, dp1 = p2 - p0;
  SurfaceInteraction si{};
  si.P = p0 * bary.x + (p1 * bary.y + (p2 * bary.z));
  si.T = rec.T;
  si.N = Normalized(Cross(dp0, dp1));
  si.shape = this;
  if (uv.empty()) {
    si.UV = primUV;
    auto st = CoordinateSystem(si.N);
    si.dPdU = st.first;
    si.dPdV = st.second;
  } else {
    V2 uv0 = uv[f0], uv1 = uv[f1], uv2 = uv[f2];
    si.UV = V2(uv0.x * bary.x + (uv1.x * bary.y + (uv2.x * bary.z)), uv0.y * bary.x + (uv1.y * bary.y + (uv2.y * bary.z)));
    V2 duv0(uv1.x - uv0.x, uv1.y - uv0.y), duv1(uv2.x - uv0.x, uv2.y - uv0.y);
    Float det = duv0.x * duv1.y - (duv0.y * duv1.x);
    Float invDet = det == 0 ? 0 : Rcp(det);
    if (invDet == 0) {
      si.dPdU = V3();
      si.dPdV = V3();
    } else {
      si.dPdU = (duv1.y * dp0 - (duv0.y * dp1)) * invDet;
      si.dPdV = (-duv1.x * dp0 + (duv0.x * dp1)) * invDet;
    }
  }
  if (normal.empty()) {
    si.Shading.N = si.N;
  } else {
    V3 n0 = normal[f0], n1 = normal[f1], n2 = normal[f2];
    V3 shN = n0 * bary.x + (n1 * bary.y + (n2 * bary.z));
    Float il = Rsqrt(SquaredNorm(shN));
    shN *= il;
    si.Shading.N = shN;
  }
  return si;
}

// mesh_base.cpp:84-112
PositionSampleResult MeshShape::SamplePosition(const V2& xi) const {
  V2 txi = xi;
  auto [index, ry] = dist.SampleReuse(txi.y);
  txi.y = ry;
  uint32_t face = uint32_t(index) * 3;
  uint32_t f0 = indices[face + 0], f1 = indices[face + 1], f2 = indices[face + 2];
  V3 p0 = position[f0], p1 = position[f1], p2 = position[f2];
  V3 e0 = p1 - p0, e1 = p2 - p0;
  V2 b = Warp::SquareToUniformTriangle(txi);
  PositionSampleResult psr{};
  psr.P = e0 * b.x + (e1 * b.y + p0);
  psr.Pdf = dist.normalization;
  psr.IsDelta = false;
  if (uv.empty()) {
    psr.UV = b;
  } else {
    V2 uv0 = uv[f0], uv1 = uv[f1], uv2 = uv[f2];
    Float w = 1 - b.x - b.y;
    psr.UV = V2(uv0.x * w + (uv1.x * b.x + (uv2.x * b.y)), uv0.y * w + (uv1.y * b.x + (uv2.y * b.y)));
  }
  if (normal.empty()) {
    psr.N = Normalized(Cross(e0, e1));
  } else {
    V3 n0 = normal[f0], n1 = normal[f1], n2 = normal[f2];
    psr.N = Normalized(n0 * (1 - b.x - b.y) + (n1 * b.x + (n2 * b.y)));
  }
  return psr;
}

// mesh_base.cpp:120-131
void MeshShape::UpdateDistribution() {
  std::vector<Float> area;
  area.reserve(triangleCount);
  for (uint32_t i = 0; i < triangleCount; i++) {
    V3 p0 = position[indices[i * 3 + 0]], p1 = position[indices[i * 3 + 1]], p2 = position[indices[i * 3 + 2]];
    area.push_back(Norm(Cross(p1 - p0, p2 - p0)) * Float(0.5));
  }
  dist.Build(area);
  surfaceArea = dist.sum;
}

// shape.cpp:7-25
DirectionSampleResult Shape::SampleDirection(const Interaction& ref, const V2& xi) const {
  PositionSampleResult psr = SamplePosition(xi);
  DirectionSampleResult dsr{};
  static_cast<PositionSampleResult&>(dsr) = psr;
  dsr.Dir = dsr.P - ref.P;
  Float distanceSquared = SquaredNorm(dsr.Dir);
  dsr.Dist = std::sqrt(distanceSquared);
  dsr.Dir /= dsr.Dist;
  Float dp = std::abs(Dot(dsr.Dir, dsr.N));
  Float x = distanceSquared / dp;
  dsr.Pdf *= std::isfinite(x) ? x : 0;
  return dsr;
}

Float Shape::PdfDirection(const Interaction& ref, const DirectionSampleResult& dsr) const {
  Float pdf = PdfPosition(dsr);
  Float dp = std::abs(Dot(dsr.Dir, dsr.N));
  pdf *= dp == 0 ? 0 : Sqr(dsr.Dist) / dp;
  return pdf;
}

// ---------------------------------------------------------------- sphere
static std::tuple<bool, Float, Float> SolveQuadratic(Float a, Float b, Float c) {  // math_ext.h:79-97
  bool linearCase = a == 0.0f;
  bool validLinear = linearCase && b != 0.0f;
  Float x0 = -c / b, x1 = -c / b;
  Float discrim = std::fma(b, b, -((Float(4) * a) * c));
  bool validQuadratic = !linearCase && (discrim >= 0.0f);
  if (validQuadratic) {
    Float rootDiscrim = std::sqrt(discrim);
    Float temp = Float(-0.5) * (b + std::copysign(rootDiscrim, b));
    Float x0p = temp / a;
    Float x1p = c / temp;
    Float x0m = std::min(x0p, x1p);
    Float x1m = std::max(x0p, x1p);
    x0 = linearCase ? x0 : x0m;
    x1 = linearCase ? x0 : x1m;
  }
  return {validLinear || validQuadratic, x0, x1};
}

Bounds SphereShape::PrimBounds(uint32_t) const {
  Bounds b;
  b.Extend(center - V3(radius));
  b.Extend(center + V3(radius));
  return b;
}

// sphere.cpp:40-54, 71-120
bool SphereShape::IntersectPrim(uint32_t, const Ray& ray, Float near, Float far, HitShapeRecord& rec) const {
  V3 o = ray.O - center;
  Float a = DotX(ray.D, ray.D);
  Float b = 2 * DotX(o, ray.D);
  Float c = std::fma(-radius, radius, DotX(o, o));
  auto [found, nearT, farT] = SolveQuadratic(a, b, c);
  bool outBounds = !(nearT <= far && farT >= near);
  bool inBounds = nearT < near && farT > far;
  bool isHit = found && !outBounds && !inBounds;
  if (!isHit) return false;
  Float t = nearT < 0 ? farT : nearT;
  rec.T = t;
  rec.PrimitiveUV = V2(0, 0);
  rec.PrimitiveIndex = 0;
  rec.ShapeIndex = index;
  rec.GeometryNormal = Normalized((ray.D * t + ray.O) - center);
  return true;
}

// UnitAngleZ (math_ext.h:101-104)
static Float UnitAngleZ(const V3& v) {
  Float temp = std::asin(Float(0.5) * Norm(V3(v.x, v.y, v.z - MulSign(Float(1), v.z)))) * 2;
  return v.z >= 0 ? temp : PI - temp;
}

// sphere.cpp:215-249
SurfaceInteraction SphereShape::ComputeInteraction(const Ray&, const HitShapeRecord& rec) const {
  SurfaceInteraction si{};
  si.T = rec.T;
  si.Shading.N = rec.GeometryNormal;
  si.P = Fmadd(si.Shading.N, V3(radius), center);
  V3 local = toWorld.ApplyAffineToLocal(si.P);
  Float rd2 = Sqr(local.x) + Sqr(local.y);
  Float theta = UnitAngleZ(local);
  Float phi = std::atan2(local.y, local.x);
  if (phi < 0) phi += 2 * PI;
  si.UV = V2(phi / (2 * PI), theta / PI);
  si.dPdU = V3(-local.y, local.x, 0.f);
  Float rd = std::sqrt(rd2);
  Float invRd = Rcp(rd);
  Float cosPhi = local.x * invRd;
  Float sinPhi = local.y * invRd;
  si.dPdV = V3(local.z * cosPhi, local.z * sinPhi, -rd);
  if (rd == 0) si.dPdV = V3(1, 0, 0);
  si.dPdU = toWorld.ApplyLinearToWorld(si.dPdU * (2 * PI));
  si.dPdV = toWorld.ApplyLinearToWorld(si.dPdV * PI);
  si.N = si.Shading.N;
  si.shape = this;
  return si;
}

PositionSampleResult SphereShape::SamplePosition(const V2& xi) const {  // sphere.cpp:251-259
  V3 local = Warp::SquareToUniformSphere(xi);
  PositionSampleResult psr{};
  psr.P = Fmadd(local, V3(radius), center);
  psr.N = local;
  psr.IsDelta = radius == 0;
  psr.Pdf = 1 / surfaceArea;
  return psr;
}

// sphere.cpp:265-320
DirectionSampleResult SphereShape::SampleDirection(const Interaction& ref, const V2& xi) const {
  V3 dcv = center - ref.P;
  Float dc2 = SquaredNorm(dcv);
  Float radiusAdj = radius * (Float(1) - RayEpsilon);
  bool isOut = dc2 > Sqr(radiusAdj);
  DirectionSampleResult dsr{};
  if (isOut) {
    Float invDc = Rsqrt(dc2);
    Float sinThetaMax = radius * invDc;
    Float sinThetaMax2 = Sqr(sinThetaMax);
    Float invSinThetaMax = Rcp(sinThetaMax);
    Float cosThetaMax = SafeSqrt(1.f - sinThetaMax2);
    Float sinTheta2 = sinThetaMax2 > Float(0.00068523262) ? Float(1) - Sqr(Fmadd(cosThetaMax - Float(1), xi.x, Float(1)))
                                                          : sinThetaMax2 * xi.x;
    Float cosTheta = SafeSqrt(Float(1) - sinTheta2);
    Float cosAlpha = sinTheta2 * invSinThetaMax + cosTheta * SafeSqrt(Fmadd(-sinTheta2, Sqr(invSinThetaMax), Float(1)));
    Float sinAlpha = SafeSqrt(Fmadd(-cosAlpha, cosAlpha, Float(1)));
    Float ang = xi.y * (Float(2) * PI);
    Float sinPhi = std::sin(ang), cosPhi = std::cos(ang);
    V3 d = Frame(dcv * -invDc).ToWorld(V3(cosPhi * sinAlpha, sinPhi * sinAlpha, cosAlpha));
    dsr.P = Fmadd(d, V3(radius), center);
    dsr.N = d;
    dsr.Dir = dsr.P - ref.P;
    Float dist2 = SquaredNorm(dsr.Dir);
    dsr.Dist = std::sqrt(dist2);
    dsr.Dir = dsr.Dir / dsr.Dist;
    dsr.Pdf = Warp::SquareToUniformConePdf(cosThetaMax);
  } else {
    V3 d = Warp::SquareToUniformSphere(xi);
    dsr.P = Fmadd(d, V3(radius), center);
    dsr.N = d;
    dsr.Dir = dsr.P - ref.P;
    Float dist2 = SquaredNorm(dsr.Dir);
    dsr.Dist = std::sqrt(dist2);
    dsr.Dir = dsr.Dir / dsr.Dist;
    dsr.Pdf = (1 / surfaceArea) * dist2 / AbsDot(dsr.Dir, dsr.N);
  }
  dsr.IsDelta = radius == 0;
  return dsr;
}

Float SphereShape::PdfDirection(const Interaction& ref, const DirectionSampleResult& dsr) const {  // sphere.cpp:322-328
  Float sinAlpha = radius * Rcp(Norm(center - ref.P));
  Float cosAlpha = SafeSqrt(Float(1) - sinAlpha * sinAlpha);
  return sinAlpha < OneMinusEpsilon ? Warp::SquareToUniformConePdf(cosAlpha)
                                    : (1 / surfaceArea) * Sqr(dsr.Dist) / AbsDot(dsr.Dir, dsr.N);
}

// ---------------------------------------------------------------- rectangle
Bounds RectangleShape::PrimBounds(uint32_t) const {  // rectangle.cpp:48-64
  Bounds b;
  b.Extend(toWorld.ApplyAffineToWorld(V3(-1, -1, 0)));
  b.Extend(toWorld.ApplyAffineToWorld(V3(1, -1, 0)));
  b.Extend(toWorld.ApplyAffineToWorld(V3(1, 1, 0)));
  b.Extend(toWorld.ApplyAffineToWorld(V3(-1, 1, 0)));
  return b;
}

// rectangle.cpp:34-46: ray taken to object space with an unnormalised direction.
bool RectangleShape::IntersectPrim(uint32_t, const Ray& ray, Float near, Float far, HitShapeRecord& rec) const {
  const auto& M = toObject;
  auto P = [&](int r) {
    return std::fma(M[r][0], ray.O.x, std::fma(M[r][1], ray.O.y, std::fma(M[r][2], ray.O.z, M[r][3])));
  };
  auto L = [&](int r) { return std::fma(M[r][0], ray.D.x, std::fma(M[r][1], ray.D.y, M[r][2] * ray.D.z)); };
  V3 o(P(0), P(1), P(2));
  V3 d(L(0), L(1), L(2));
  Float t = -o.z / d.z;
  V3 local(std::fma(d.x, t, o.x), std::fma(d.y, t, o.y), std::fma(d.z, t, o.z));
  bool isHit = t >= near && t <= far && std::abs(local.x) <= 1 && std::abs(local.y) <= 1;
  if (!isHit) return false;
  rec.T = t;
  rec.PrimitiveUV = V2(local.x, local.y);
  rec.PrimitiveIndex = 0;
  rec.ShapeIndex = index;
  rec.GeometryNormal = frame.N;
  return true;
}

SurfaceInteraction RectangleShape::ComputeInteraction(const Ray& ray, const HitShapeRecord& rec) const {  // :204-219
  SurfaceInteraction si{};
  si.T = rec.T;
  V3 p = ray(rec.T);
  si.P = Dot(toWorld.TranslationToWorld() - p, frame.N) * frame.N + p;
  si.N = frame.N;
  si.Shading.N = frame.N;
  si.dPdU = frame.S;
  si.dPdV = frame.T;
  si.UV = V2(rec.PrimitiveUV.x * Float(0.5) + Float(0.5), rec.PrimitiveUV.y * Float(0.5) + Float(0.5));
  si.shape = this;
  return si;
}

PositionSampleResult RectangleShape::SamplePosition(const V2& xi) const {  // :221-230
  PositionSampleResult psr{};
  V2 s(xi.x * 2 - 1, xi.y * 2 - 1);
  psr.P = toWorld.ApplyAffineToWorld(V3(s.x, s.y, 0));
  psr.N = frame.N;
  psr.Pdf = 1 / surfaceArea;
  psr.UV = xi;
  psr.IsDelta = false;
  return psr;
}

// ---------------------------------------------------------------- distributions
void Dist1D::Build(const Float* p, size_t n) {  // distribution.cpp:55-68
  pdf.assign(p, p + n);
  cdf.assign(n + 1, 0);
  for (size_t i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + pdf[i - 1] / Float(n);
  sum = cdf[n];
  for (size_t i = 1; i < n + 1; i++) cdf[i] /= sum;
  normalization = Float(1.0 / sum);
}

// distribution.cpp:85-96: the cell below the first cdf entry >= xi (the
// unsigned index - 1 wraps for xi <= cdf[0] and clamps to the last cell).
std::tuple<size_t, Float, Float> Dist1D::Sample(Float xi) const {
  size_t index = size_t(std::lower_bound(cdf.begin(), cdf.end(), xi) - cdf.begin());
  index = std::clamp(index - 1, size_t(0), cdf.size() - 2);
  Float du = xi - cdf[index];
  if ((cdf[index + 1] - cdf[index]) > 0) du /= (cdf[index + 1] - cdf[index]);
  return {index, (Float(index) + du) / Float(Count()), pdf[index] * normalization};
}

void Dist2D::Build(const Float* data, size_t nu, size_t nv) {  // distribution.cpp:101-113
  conditional.resize(nv);
  for (size_t v = 0; v < nv; v++) conditional[v].Build(data + v * nu, nu);
  std::vector<Float> m(nv);
  for (size_t v = 0; v < nv; v++) m[v] = conditional[v].sum;
  marginal.Build(m.data(), nv);
}

std::tuple<size_t, size_t, Float, V2> Dist2D::Sample(const V2& xi) const {  // :115-122
  auto [vj, d1, pdf1] = marginal.Sample(xi.y);
  auto [vi, d0, pdf0] = conditional[vj].Sample(xi.x);
  return {vi, vj, pdf0 * pdf1, V2(d0, d1)};
}

Float Dist2D::Pdf(const V2& uv) const {  // :124-128
  size_t iu = std::clamp(size_t(uv.x * Float(conditional[0].Count())), size_t(0), conditional[0].Count() - 1);
  size_t iv = std::clamp(size_t(uv.y * Float(marginal.Count())), size_t(0), marginal.Count() - 1);
  return conditional[iv].pdf[iu] / marginal.sum;
}

SurfaceInteraction RectangleParamSurface(const Shape* shape, const V2& uv) {  // rectangle.cpp:236-250
  const auto* r = static_cast<const RectangleShape*>(shape);
  SurfaceInteraction si{};
  si.T = 0;
  si.P = r->toWorld.ApplyAffineToWorld(V3(uv.x * 2 - 1, uv.y * 2 - 1, 0));
  si.N = r->frame.N;
  si.Shading = r->frame;
  si.UV = uv;
  si.dPdU = r->frame.S;
  si.dPdV = r->frame.T;
  si.shape = shape;
  return si;
}

// ---------------------------------------------------------------- lights
// light/area.cpp:20-148.  A radiance texture larger than one texel is sampled
// through a 2-D distribution of its luminance at texel corners; the sample
// lands on the texel's corner uv.  SampleDirection builds its record with
// SurfaceInteraction::ToDsr, which sets the density to 0 (interaction.cpp:
// 59-75), so with quirks on such a light receives no light samples and is
// reached only by BSDF sampling (still MIS-weighted against PdfDirection).
// Quirks off: the continuous position inside the texel, with the
// PdfDirection density.
struct AreaLight final : Light {
  Spectrum radiance;
  const Texture* tex = nullptr;
  bool uniform = true;
  bool quirks = true;
  Dist2D dist;
  Spectrum Radiance(const V2& uv) const { return tex ? tex->Eval(uv, V2(), V2()) : radiance; }
  Spectrum Eval(const SurfaceInteraction& si) const override {
    return Dot(si.ToWorld(si.Wi), si.N) >= 0 ? Radiance(si.UV) : Spectrum(0);
  }
  void Init(uint32_t w, uint32_t h) {
    uniform = !tex || (w == 1 && h == 1);
    if (uniform) return;
    std::vector<Float> lum(size_t(w) * h);
    for (uint32_t y = 0; y < h; y++) {
      Float v = Float(y) / Float(h);
      for (uint32_t x = 0; x < w; x++) lum[size_t(y) * w + x] = Luminance(tex->Eval(V2(Float(x) / Float(w), v), V2(), V2()));
    }
    dist.Build(lum.data(), w, h);
  }
  std::pair<DirectionSampleResult, Spectrum> SampleDirection(const Interaction& ref, const V2& xi) const override {
    if (uniform) {
      DirectionSampleResult dsr = shape->SampleDirection(ref, xi);
      Spectrum r = Dot(-dsr.Dir, dsr.N) >= 0 ? Radiance(dsr.UV) : Spectrum(0);
      return {dsr, r};
    }
    auto [vi, vj, pdf, cont] = dist.Sample(xi);
    if (pdf <= 0) return {{}, Spectrum(0)};
    V2 uv = quirks ? V2(Float(vi) / Float(dist.conditional[0].Count()), Float(vj) / Float(dist.marginal.Count())) : cont;
    SurfaceInteraction surface = RectangleParamSurface(shape, uv);
    DirectionSampleResult dsr = surface.ToDsr(ref);
    dsr.Pdf = quirks ? 0 : PdfDirection(ref, dsr);
    Spectrum r = Dot(-dsr.Dir, dsr.N) >= 0 ? Radiance(surface.UV) : Spectrum(0);
    return {dsr, r};
  }
  Float PdfDirection(const Interaction& ref, const DirectionSampleResult& dsr) const override {
    if (uniform) return Dot(-dsr.Dir, dsr.N) >= 0 ? shape->PdfDirection(ref, dsr) : 0;
    SurfaceInteraction si = RectangleParamSurface(shape, dsr.UV);
    Float pdf = dist.Pdf(dsr.UV);
    if (pdf <= 0) return 0;
    Float dp = Dot(dsr.Dir, dsr.N);
    return pdf * Sqr(dsr.Dist) / (Norm(Cross(si.dPdU, si.dPdV)) * -dp);
  }
};

// light/skybox.cpp:24-141, 215-234: a lat-long radiance map; a map larger
// than one texel is sampled through its sin(theta)-weighted luminance at
// texel corners.  PdfDirection maps the direction with the affine inverse
// (translation included, Q15).  Quirks off: the continuous position inside
// the texel and the linear inverse.
struct SkyboxLight final : Light {
  Spectrum radiance;
  const Texture* tex = nullptr;
  bool uniform = true;
  bool quirks = true;
  Dist2D dist;
  Transform toWorld;
  V3 sphereCenter;
  Float sphereRadius = 0;
  bool IsEnv() const override { return true; }
  static V2 DirUV(const V3& v) {
    Float phi = std::atan2(v.x, -v.z);
    phi = (phi < 0) ? (phi + 2 * PI) : phi;
    Float theta = std::acos(std::clamp(v.y, Float(-1), Float(1)));
    return V2(phi / (2 * PI), theta / PI);
  }
  Spectrum Radiance(const V2& uv) const { return tex ? tex->Eval(uv, V2(), V2()) : radiance; }
  void Init(uint32_t w, uint32_t h) {
    uniform = !tex || (w == 1 && h == 1);
    if (uniform) return;
    std::vector<Float> lum(size_t(w) * h);
    for (uint32_t y = 0; y < h; y++) {
      Float sinTheta = std::sin((Float(y) + Float(0.5)) / Float(h) * PI);
      Float v = Float(y) / Float(h);
      for (uint32_t x = 0; x < w; x++) {
        Float l = Luminance(tex->Eval(V2(Float(x) / Float(w), v), V2(), V2()));
        lum[size_t(y) * w + x] = l * sinTheta;
      }
    }
    dist.Build(lum.data(), w, h);
  }
  Spectrum Eval(const SurfaceInteraction& si) const override {
    return tex ? Radiance(DirUV(toWorld.ApplyLinearToLocal(-si.Wi))) : radiance;
  }
  std::pair<DirectionSampleResult, Spectrum> SampleDirection(const Interaction& ref, const V2& xi) const override {
    Float radius = std::max(sphereRadius, Norm(ref.P - sphereCenter));
    Float dist2 = 2 * radius;
    DirectionSampleResult dsr{};
    if (uniform) {
      V3 dir = Normalized(Warp::SquareToUniformSphere(xi));
      V3 worldDir = toWorld.ApplyLinearToWorld(dir);
      dsr.P = worldDir * dist2 + ref.P;
      dsr.N = -worldDir;
      dsr.UV = DirUV(dir);
      dsr.Pdf = Warp::SquareToUniformSpherePdf();
      dsr.Dir = worldDir;
      dsr.Dist = dist2;
      dsr.IsDelta = false;
      return {dsr, Radiance(dsr.UV)};
    }
    auto [vi, vj, pdf, cont] = dist.Sample(xi);
    if (pdf <= 0) return {dsr, Spectrum(0)};
    Float y = quirks ? Float(vj) / Float(dist.marginal.Count()) : cont.y;
    Float x = quirks ? Float(vi) / Float(dist.conditional[0].Count()) : cont.x;
    Float theta = y * PI, phi = x * 2 * PI;
    Float sinTheta = std::sin(theta);
    if (sinTheta == 0) return {dsr, Spectrum(0)};
    pdf /= (2 * PI * PI * sinTheta);
    V3 sph(std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, std::cos(theta));  // math_ext.h:108-112
    V3 worldDir = toWorld.ApplyLinearToWorld(V3(sph.y, sph.z, -sph.x));
    dsr.P = worldDir * dist2 + ref.P;
    dsr.N = -worldDir;
    dsr.UV = V2(x, y);
    dsr.Pdf = pdf;
    dsr.Dir = worldDir;
    dsr.Dist = dist2;
    dsr.IsDelta = false;
    return {dsr, Radiance(dsr.UV)};
  }
  Float PdfDirection(const Interaction&, const DirectionSampleResult& dsr) const override {
    if (uniform) return Warp::SquareToUniformSpherePdf();
    V3 v = quirks ? toWorld.ApplyAffineToLocal(dsr.Dir) : toWorld.ApplyLinearToLocal(dsr.Dir);
    Float theta = std::acos(std::clamp(v.y, Float(-1), Float(1)));
    Float sinTheta = std::sin(theta);
    if (sinTheta == 0) return 0;
    return dist.Pdf(DirUV(v)) / (2 * PI * PI * sinTheta);
  }
  void SetScene(const Scene& scene) override {  // skybox.cpp:136-141, bounding_sphere.cpp:15-18
    Bounds b = scene.accel.world;
    V3 c = b.Center();
    sphereCenter = c;
    sphereRadius = std::max(RayEpsilon, Norm(c - b.hi) * (1 + ShadowEpsilon));
  }
};

struct PointLight final : Light {  // light/point.cpp:10-45
  V3 pos;
  Spectrum intensity;
  Spectrum Eval(const SurfaceInteraction&) const override { return Spectrum(0); }
  std::pair<DirectionSampleResult, Spectrum> SampleDirection(const Interaction& ref, const V2&) const override {
    DirectionSampleResult dsr{};
    dsr.P = pos;
    dsr.Pdf = 1;
    dsr.IsDelta = true;
    dsr.Dir = dsr.P - ref.P;
    Float dist2 = Dot(dsr.Dir, dsr.Dir);
    dsr.Dist = std::sqrt(dist2);
    Float invDist = Rcp(dsr.Dist);
    dsr.Dir = dsr.Dir * invDist;
    return {dsr, intensity * Sqr(invDist)};
  }
  Float PdfDirection(const Interaction&, const DirectionSampleResult&) const override { return 0; }
};

// light/projection.cpp:12-63.  The irradiance is a constant (stored already
// multiplied by the scale) or a texture; a texture's width / height sets the
// aspect of the projected image (a chessboard counts as 1 x 1).
struct ProjectionLight final : Light {
  Transform toWorld, toClip;
  Spectrum irradiance;  // constant: irradiance x scale; textured: the scale
  const Texture* tex = nullptr;
  void Init(const float* tw, Float fov, uint32_t texW, uint32_t texH) {
    toWorld = Transform(Mat4::FromRowMajor(tw));
    Float aspect = Float(texW) / Float(texH), farC = 10000, nearC = Float(0.0001);
    Float recip = 1 / (farC - nearC);
    Float cot = 1 / std::tan(Radian(fov / 2));
    Float P[16] = {cot, 0, 0, 0, 0, cot, 0, 0, 0, 0, farC * recip, -nearC * farC * recip, 0, 0, 1, 0};
    Float S[16] = {-0.5f, 0, 0, 0, 0, -0.5f * aspect, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    Float T[16] = {1, 0, 0, -1, 0, 1, 0, -1 / aspect, 0, 0, 1, 0, 0, 0, 0, 1};
    toClip = Transform(Mat4::FromRowMajor(S) * Mat4::FromRowMajor(T) * Mat4::FromRowMajor(P));
  }
  Spectrum Eval(const SurfaceInteraction&) const override { return Spectrum(0); }
  std::pair<DirectionSampleResult, Spectrum> SampleDirection(const Interaction& ref, const V2&) const override {
    V3 local = toWorld.ApplyAffineToLocal(ref.P);
    V3 c = toClip.ApplyAffineToWorld(local);
    bool inRange = c.x >= 0 && c.x <= 1 && c.y >= 0 && c.y <= 1 && local.z > 0;
    if (!inRange) return {DirectionSampleResult{}, Spectrum(0)};
    DirectionSampleResult dsr{};
    dsr.P = V3(toWorld.ToWorld.m[0][3], toWorld.ToWorld.m[1][3], toWorld.ToWorld.m[2][3]);
    dsr.N = Normalized(toWorld.ApplyLinearToWorld(V3(0, 0, 1)));
    dsr.UV = V2(c.x, c.y);
    dsr.Pdf = 1;
    dsr.IsDelta = true;
    dsr.Dir = dsr.P - ref.P;
    Float sq = Dot(dsr.Dir, dsr.Dir);
    dsr.Dist = std::sqrt(sq);
    dsr.Dir = dsr.Dir * Rcp(dsr.Dist);
    // SurfaceInteraction(dsr) carries the uv and no differentials (interaction.cpp:32-43)
    Spectrum irr = tex ? Mul(tex->Eval(dsr.UV, V2(), V2()), irradiance) : irradiance;
    return {dsr, irr * PI / (Sqr(local.z) * Dot(-dsr.Dir, dsr.N))};
  }
  Float PdfDirection(const Interaction&, const DirectionSampleResult&) const override { return 0; }
};

// ---------------------------------------------------------------- camera
void Camera::Init(const rad_camera_desc& d) {  // perspective.cpp:15-70
  nearClip = d.near_clip;
  farClip = d.far_clip;
  width = d.width;
  height = d.height;
  rcpResolution = V2(1.f / width, 1.f / height);
  cameraToWorld = Transform(Mat4::FromRowMajor(d.to_world));
  Float aspect = width / static_cast<Float>(height);
  Float recip = 1 / (farClip - nearClip);
  Float cot = 1 / std::tan(Radian(d.fov / 2));
  Mat4 perspective, scale, translate;
  Float P[16] = {cot, 0, 0, 0, 0, cot, 0, 0, 0, 0, farClip * recip, -nearClip * farClip * recip, 0, 0, 1, 0};
  Float S[16] = {-0.5f, 0, 0, 0, 0, -0.5f * aspect, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  Float T[16] = {1, 0, 0, -1, 0, 1, 0, -1 / aspect, 0, 0, 1, 0, 0, 0, 0, 1};
  perspective = Mat4::FromRowMajor(P);
  scale = Mat4::FromRowMajor(S);
  translate = Mat4::FromRowMajor(T);
  cameraToClip = Transform(scale * translate * perspective);
  // perspective.cpp:55-58
  V3 c0 = cameraToClip.ApplyAffineToLocal(V3(0, 0, 0));
  dx = cameraToClip.ApplyAffineToLocal(V3(Float(1) / Float(width), 0, 0)) - c0;
  dy = cameraToClip.ApplyAffineToLocal(V3(0, Float(1) / Float(height), 0)) - c0;
}

Ray Camera::SampleRay(const V2& sp) const {  // perspective.cpp:84-99 (differentials unused)
  V3 ndc(sp.x * rcpResolution.x, sp.y * rcpResolution.y, 0);
  V3 nearP = cameraToClip.ApplyAffineToLocal(ndc);
  V3 dir = Normalized(nearP);
  Float invZ = Rcp(dir.z);
  V3 o = cameraToWorld.ApplyAffineToWorld(V3(0, 0, 0));
  V3 d = Normalized(cameraToWorld.ApplyLinearToWorld(dir));
  return Ray{o, d, nearClip * invZ, farClip * invZ};
}

Ray Camera::SampleRayDifferential(const V2& sp, RayDiff& rd) const {  // perspective.cpp:84-99
  V3 ndc(sp.x * rcpResolution.x, sp.y * rcpResolution.y, 0);
  V3 nearP = cameraToClip.ApplyAffineToLocal(ndc);
  Ray r = SampleRay(sp);
  rd.has = true;
  rd.Ox = r.O;
  rd.Oy = r.O;
  rd.Dx = cameraToWorld.ApplyLinearToWorld(Normalized(nearP + dx));
  rd.Dy = cameraToWorld.ApplyLinearToWorld(Normalized(nearP + dy));
  return r;
}

// ---------------------------------------------------------------- scene
const Medium* SurfaceInteraction::GetMedium(const Ray& ray) const {  // interaction.cpp:113-115
  return Dot(ray.D, N) > 0 ? shape->outside : shape->inside;
}

static const float kCubeV[24][3] = {{1, -1, -1}, {1, -1, 1}, {-1, -1, 1}, {-1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                    {-1, 1, 1},  {1, 1, 1},  {1, -1, -1}, {1, 1, -1},   {1, 1, 1},   {1, -1, 1},
                                    {1, -1, 1},  {1, 1, 1},  {-1, 1, 1},  {-1, -1, 1},  {-1, -1, 1}, {-1, 1, 1},
                                    {-1, 1, -1}, {-1, -1, -1}, {1, 1, -1}, {1, -1, -1}, {-1, -1, -1}, {-1, 1, -1}};
static const float kCubeN[6][3] = {{0, -1, 0}, {0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {-1, 0, 0}, {0, 0, -1}};
static const uint32_t kCubeT[12][3] = {{0, 1, 2},    {3, 0, 2},    {4, 5, 6},    {7, 4, 6},    {8, 9, 10},   {11, 8, 10},
                                       {12, 13, 14}, {15, 12, 14}, {16, 17, 18}, {19, 16, 18}, {20, 21, 22}, {23, 20, 22}};

void Scene::Create(const rad_scene_desc& desc, bool q) {
  quirks = q;
  BuildTextures(desc.textures, desc.texture_count, textures);
  bsdfDescs.assign(desc.bsdfs, desc.bsdfs + desc.bsdf_count);
  specWeights.assign(desc.bsdf_count, std::numeric_limits<Float>::quiet_NaN());
  bsdfTextured.assign(desc.bsdf_count, false);
  for (uint32_t i = 0; i < desc.bsdf_count; i++) {
    const rad_bsdf_desc& d = desc.bsdfs[i];
    for (int k = 0; k < RAD_TEX_SLOTS; k++)
      if (d.tex[k] < 0 || uint32_t(d.tex[k]) > desc.texture_count) throw std::runtime_error("bsdf: texture index out of range");
    // plastic.cpp:29-42 / rough_plastic.cpp:42-55: the sampling weight from the texel means
    if ((d.type == RAD_BSDF_PLASTIC || d.type == RAD_BSDF_ROUGH_PLASTIC) &&
        (d.tex[RAD_TEX_DIFFUSE] || d.tex[RAD_TEX_SPECULAR])) {
      double dm = d.tex[RAD_TEX_DIFFUSE] ? textures[d.tex[RAD_TEX_DIFFUSE] - 1]->MeanLuminance()
                                         : double(Luminance(Spectrum(d.diffuse[0], d.diffuse[1], d.diffuse[2])));
      double sm = d.tex[RAD_TEX_SPECULAR]
                      ? textures[d.tex[RAD_TEX_SPECULAR] - 1]->MeanLuminance()
                      : double(Luminance(Spectrum(d.specular_reflectance[0], d.specular_reflectance[1], d.specular_reflectance[2])));
      specWeights[i] = Float(sm / (dm + sm));
    }
  }
  for (uint32_t i = 0; i < desc.bsdf_count; i++) {
    for (int32_t j = int32_t(i), guard = 0; j >= 0 && uint32_t(j) < desc.bsdf_count && guard < 3; guard++) {
      const rad_bsdf_desc& d = desc.bsdfs[j];
      for (int k = 0; k < RAD_TEX_SLOTS; k++) bsdfTextured[i] = bsdfTextured[i] || d.tex[k] != 0;
      j = (d.type == RAD_BSDF_TWO_SIDE || d.type == RAD_BSDF_MASK) ? d.nested : -1;
    }
  }
  for (uint32_t i = 0; i < desc.bsdf_count; i++)
    bsdfs.push_back(MakeBsdf(desc.bsdfs, desc.bsdf_count, i, quirks, 0, specWeights.data()));
  for (uint32_t i = 0; i < desc.medium_count; i++) media.push_back(MakeMedium(desc.media[i]));
  globalMedium = desc.global_medium >= 0 ? media[desc.global_medium].get() : nullptr;
  for (uint32_t i = 0; i < desc.light_count; i++) {
    const rad_light_desc& ld = desc.lights[i];
    auto texOf = [&](int32_t k) -> const Texture* {
      if (k == 0) return nullptr;
      if (k < 0 || uint32_t(k) > desc.texture_count || desc.textures[k - 1].channels != 3)
        throw std::runtime_error("light: bad radiance texture");
      return textures[size_t(k - 1)].get();
    };
    auto texSize = [&](int32_t k, uint32_t& w, uint32_t& h) {
      w = h = 1;
      if (k > 0 && desc.textures[k - 1].type == RAD_TEXTURE_BITMAP) w = desc.textures[k - 1].width, h = desc.textures[k - 1].height;
    };
    uint32_t tw = 1, th = 1;
    texSize(ld.tex, tw, th);
    if (ld.type == RAD_LIGHT_AREA) {
      auto l = std::make_unique<AreaLight>();
      l->radiance = Spectrum(ld.radiance[0], ld.radiance[1], ld.radiance[2]);
      l->tex = texOf(ld.tex);
      l->quirks = quirks;
      l->Init(tw, th);
      lights.push_back(std::move(l));
    } else if (ld.type == RAD_LIGHT_POINT) {
      auto l = std::make_unique<PointLight>();
      l->intensity = Spectrum(ld.radiance[0], ld.radiance[1], ld.radiance[2]);
      l->pos = V3(ld.to_world[3], ld.to_world[7], ld.to_world[11]);
      lights.push_back(std::move(l));
    } else if (ld.type == RAD_LIGHT_PROJECTION) {
      auto l = std::make_unique<ProjectionLight>();
      l->irradiance = Spectrum(ld.radiance[0], ld.radiance[1], ld.radiance[2]);
      l->tex = texOf(ld.tex);
      l->Init(ld.to_world, ld.fov, tw, th);
      lights.push_back(std::move(l));
    } else {
      auto l = std::make_unique<SkyboxLight>();
      l->radiance = Spectrum(ld.radiance[0], ld.radiance[1], ld.radiance[2]);
      l->toWorld = Transform(Mat4::FromRowMajor(ld.to_world));
      l->tex = texOf(ld.tex);
      l->quirks = quirks;
      l->Init(tw, th);
      lights.push_back(std::move(l));
    }
  }
  for (uint32_t i = 0; i < desc.shape_count; i++) {
    const rad_shape_desc& sd = desc.shapes[i];
    Mat4 tw = Mat4::FromRowMajor(sd.to_world);
    std::unique_ptr<Shape> s;
    if (sd.type == RAD_SHAPE_MESH || sd.type == RAD_SHAPE_CUBE) {
      auto m = std::make_unique<MeshShape>();
      m->toWorld = Transform(tw);
      bool ident = m->toWorld.IsIdentity();
      if (sd.type == RAD_SHAPE_MESH) {
        m->triangleCount = sd.triangle_count;
        m->indices.assign(sd.indices, sd.indices + 3 * size_t(sd.triangle_count));
        m->position.resize(sd.vertex_count);
        for (uint32_t v = 0; v < sd.vertex_count; v++) {
          V3 p(sd.positions[3 * v], sd.positions[3 * v + 1], sd.positions[3 * v + 2]);
          m->position[v] = ident ? p : m->toWorld.ApplyAffineToWorld(p);
        }
        if (sd.normals) {
          m->normal.resize(sd.vertex_count);
          for (uint32_t v = 0; v < sd.vertex_count; v++) {
            V3 n(sd.normals[3 * v], sd.normals[3 * v + 1], sd.normals[3 * v + 2]);
            m->normal[v] = ident ? n : m->toWorld.ApplyNormalToWorld(n);
          }
        }
        if (sd.uvs) {
          m->uv.resize(sd.vertex_count);
          for (uint32_t v = 0; v < sd.vertex_count; v++) m->uv[v] = V2(sd.uvs[2 * v], sd.uvs[2 * v + 1]);
        }
      } else {  // cube.cpp:15-123; UVs are never initialised there, taken as zero here
        m->triangleCount = 12;
        m->position.resize(24);
        m->normal.resize(24);
        m->uv.assign(24, V2(0, 0));
        for (int v = 0; v < 24; v++) {
          m->position[v] = m->toWorld.ApplyAffineToWorld(V3(kCubeV[v][0], kCubeV[v][1], kCubeV[v][2]));
          const float* n = kCubeN[v / 4];
          m->normal[v] = m->toWorld.ApplyNormalToWorld(V3(n[0], n[1], n[2]));
        }
        for (int t = 0; t < 12; t++)
          for (int k = 0; k < 3; k++) m->indices.push_back(kCubeT[t][k]);
      }
      m->UpdateDistribution();
      s = std::move(m);
    } else if (sd.type == RAD_SHAPE_SPHERE) {  // sphere.cpp:173-190
      auto sp = std::make_unique<SphereShape>();
      Transform transform(tw);
      if (transform.HasNonUniformScale()) throw std::runtime_error("Sphere does not support non uniform transformations");
      sp->center = transform.ApplyAffineToWorld(V3(sd.center[0], sd.center[1], sd.center[2]));
      sp->radius = Norm(transform.ApplyLinearToWorld(V3(0, 0, sd.radius)));
      Float TS[16] = {sp->radius, 0, 0, sp->center.x, 0, sp->radius, 0, sp->center.y, 0, 0, sp->radius, sp->center.z, 0, 0, 0, 1};
      sp->toWorld = Transform(Mat4::FromRowMajor(TS) * tw);
      sp->surfaceArea = 4 * PI * Sqr(sp->radius);
      s = std::move(sp);
    } else if (sd.type == RAD_SHAPE_RECTANGLE) {  // rectangle.cpp:164-179
      auto r = std::make_unique<RectangleShape>();
      if (sd.flip_normal) {
        Float F[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1};
        r->toWorld = Transform(tw * Mat4::FromRowMajor(F));
      } else {
        r->toWorld = Transform(tw);
      }
      V3 dpdu = r->toWorld.ApplyLinearToWorld(V3(2, 0, 0));
      V3 dpdv = r->toWorld.ApplyLinearToWorld(V3(0, 2, 0));
      V3 n = Normalized(r->toWorld.ApplyNormalToWorld(V3(0, 0, 1)));
      r->frame = Frame(dpdu, dpdv, n);
      r->surfaceArea = Norm(Cross(dpdu, dpdv));
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) r->toObject[a][b] = r->toWorld.ToLocal.m[a][b];
      s = std::move(r);
    } else {
      throw std::runtime_error("unknown shape type");
    }
    s->index = i;
    s->bsdfIndex = sd.bsdf;
    if (sd.bsdf >= 0) s->bsdf = bsdfs[sd.bsdf].get();
    if (sd.light >= 0) {
      s->light = lights[sd.light].get();
      lights[sd.light]->shape = s.get();
      auto* al = dynamic_cast<AreaLight*>(lights[sd.light].get());
      if (al && !al->uniform && sd.type != RAD_SHAPE_RECTANGLE)
        throw std::runtime_error("textured area light: only rectangles have a parametric surface (shape.h EvalParamSurface)");
    }
    s->inside = sd.medium_inside >= 0 ? media[sd.medium_inside].get() : nullptr;
    s->outside = sd.medium_outside >= 0 ? media[sd.medium_outside].get() : nullptr;
    shapes.push_back(std::move(s));
  }
  camera.Init(desc.camera);
  accel.Build(shapes);
  lightPdf = Rcp(Float(lights.size()));
  for (auto& l : lights) {
    if (l->IsEnv()) {
      envLight = l.get();
      l->SetScene(*this);
    }
  }
}

bool Scene::RayIntersect(const Ray& ray, SurfaceInteraction& si) const {  // scene.cpp:42-51
  HitShapeRecord rec;
  bool hit = brute ? accel.ClosestBrute(ray, rec) : accel.Closest(ray, rec);
  if (!hit) {
    si = {};
    si.Wi = -ray.D;
    return false;
  }
  si = shapes[rec.ShapeIndex]->ComputeInteraction(ray, rec);  // interaction.cpp:121-127
  si.ComputeShadingFrame();
  si.Wi = si.ToLocal(-ray.D);
  si.ShapeIndex = rec.ShapeIndex;
  return true;
}

bool Scene::IsOcclude(const Interaction& ref, const V3& p) const {  // scene.cpp:53-63
  V3 o = ref.OffsetP(p - ref.P);
  V3 d = p - o;
  Float dist = Norm(d);
  if (dist < ShadowEpsilon) return true;
  d /= dist;
  Ray shadowRay{o, d, 0, dist * (1 - ShadowEpsilon)};
  return RayIntersect(shadowRay);
}

std::pair<uint32_t, Float> Scene::SampleLight(Float xi) const {  // scene.cpp:65-77
  if (lights.size() < 2) {
    if (lights.size() == 1) return {0u, Float(1)};
    return {0xffffffffu, Float(0)};
  }
  uint32_t lightCount = (uint32_t)lights.size();
  Float lightIndex = xi * lightCount;
  uint32_t index = std::min(uint32_t(lightIndex), lightCount - 1);
  return {index, lightPdf};
}

std::tuple<const Light*, DirectionSampleResult, Spectrum> Scene::SampleLightDirection(const Interaction& ref, Float sl,
                                                                                       const V2& xi) const {
  auto [index, pdf] = SampleLight(sl);
  const Light* light = lights[index].get();
  auto [dsr, li] = light->SampleDirection(ref, xi);
  dsr.Pdf *= pdf;
  return {light, dsr, li};
}

// ---------------------------------------------------------------- accel
int Accel::BuildRecursive(std::vector<PrimRef>& refs, std::vector<Bounds>& pb, std::vector<V3>& pc, uint32_t begin,
                          uint32_t end, int depth) {
  Node node;
  for (uint32_t i = begin; i < end; i++) node.b.Extend(pb[i]);
  uint32_t n = end - begin;
  int idx;
  {
    idx = (int)nodes.size();
    nodes.push_back(node);
  }
  if (n <= 4 || depth > 60) {
    nodes[idx].start = begin;
    nodes[idx].count = n;
    return idx;
  }
  Bounds cb;
  for (uint32_t i = begin; i < end; i++) cb.Extend(pc[i]);
  int bestAxis = -1, bestSplit = -1;
  Float bestCost = Inf;
  constexpr int B = 16;
  for (int a = 0; a < 3; a++) {
    Float lo = cb.lo[a], hi = cb.hi[a];
    if (!(hi > lo)) continue;
    Bounds bb[B];
    int cnt[B] = {0};
    Float scale = B / (hi - lo);
    for (uint32_t i = begin; i < end; i++) {
      int k = std::min(B - 1, int((pc[i][a] - lo) * scale));
      bb[k].Extend(pb[i]);
      cnt[k]++;
    }
    Bounds left[B];
    int lc[B];
    Bounds acc;
    int c = 0;
    for (int k = 0; k < B; k++) {
      acc.Extend(bb[k]);
      c += cnt[k];
      left[k] = acc;
      lc[k] = c;
    }
    acc = Bounds();
    c = 0;
    for (int k = B - 1; k > 0; k--) {
      acc.Extend(bb[k]);
      c += cnt[k];
      Float cost = left[k - 1].Area() * lc[k - 1] + acc.Area() * c;
      if (lc[k - 1] > 0 && c > 0 && cost < bestCost) {
        bestCost = cost;
        bestAxis = a;
        bestSplit = k;
      }
    }
  }
  uint32_t mid;
  if (bestAxis < 0) {
    mid = begin + n / 2;
  } else {
    Float lo = cb.lo[bestAxis];
    Float scale = B / (cb.hi[bestAxis] - lo);
    uint32_t i = begin, j = end;
    while (i < j) {
      int k = std::min(B - 1, int((pc[i][bestAxis] - lo) * scale));
      if (k < bestSplit) {
        i++;
      } else {
        j--;
        std::swap(refs[i], refs[j]);
        std::swap(pb[i], pb[j]);
        std::swap(pc[i], pc[j]);
      }
    }
    mid = i;
    if (mid == begin || mid == end) mid = begin + n / 2;
  }
  int l = BuildRecursive(refs, pb, pc, begin, mid, depth + 1);
  int r = BuildRecursive(refs, pb, pc, mid, end, depth + 1);
  nodes[idx].left = l;
  nodes[idx].right = r;
  return idx;
}

void Accel::Build(const std::vector<std::unique_ptr<Shape>>& s) {
  auto t0 = std::chrono::steady_clock::now();
  shapes = &s;
  std::vector<PrimRef> refs;
  std::vector<Bounds> pb;
  std::vector<V3> pc;
  for (auto& sh : s) {
    for (uint32_t p = 0; p < sh->PrimCount(); p++) {
      refs.push_back({sh->index, p});
      Bounds b = sh->PrimBounds(p);
      pb.push_back(b);
      pc.push_back(b.Center());
      world.Extend(b);
    }
  }
  nodes.clear();
  nodes.reserve(2 * refs.size() / 2 + 16);
  if (!refs.empty()) BuildRecursive(refs, pb, pc, 0, (uint32_t)refs.size(), 0);
  prims = std::move(refs);
  nodes4.clear();
  if (!nodes.empty()) Collapse(0);
  triVerts.assign(9 * prims.size(), 0.f);
  isTri.assign(prims.size(), 0);
  for (size_t i = 0; i < prims.size(); i++) {
    const auto* mesh = dynamic_cast<const MeshShape*>(s[prims[i].shape].get());
    if (!mesh) continue;
    isTri[i] = 1;
    for (int k = 0; k < 3; k++) {
      const V3& p = mesh->position[mesh->indices[prims[i].prim * 3 + k]];
      triVerts[9 * i + 3 * k] = p.x;
      triVerts[9 * i + 3 * k + 1] = p.y;
      triVerts[9 * i + 3 * k + 2] = p.z;
    }
  }
  chunks.clear();
  for (uint32_t c = 0; c < prims.size(); c += kCullChunk) {
    Bounds b;
    for (uint32_t i = c; i < std::min<uint32_t>(c + kCullChunk, (uint32_t)prims.size()); i++)
      b.Extend(s[prims[i].shape]->PrimBounds(prims[i].prim));
    chunks.push_back(b);
  }
  buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Collapse the binary subtree at `b` into 4-wide nodes: open the inner child
// of largest surface area until four children; returns the node index.
int Accel::Collapse(int b) {
  int kids[4];
  int nk = 0;
  if (nodes[b].count) {
    kids[nk++] = b;
  } else {
    kids[nk++] = nodes[b].left;
    kids[nk++] = nodes[b].right;
  }
  while (nk < 4) {
    int best = -1;
    Float bestArea = -1;
    for (int k = 0; k < nk; k++) {
      const Node& c = nodes[kids[k]];
      if (c.count) continue;
      Float a = c.b.Area();
      if (a > bestArea) {
        bestArea = a;
        best = k;
      }
    }
    if (best < 0) break;
    const int open = kids[best];
    kids[best] = nodes[open].left;
    kids[nk++] = nodes[open].right;
  }
  const int idx = (int)nodes4.size();
  nodes4.emplace_back();
  Node4 q;
  for (int k = 0; k < 4; k++) {
    if (k >= nk) {
      for (int a = 0; a < 3; a++) {
        q.lo[a][k] = Inf;
        q.hi[a][k] = -Inf;
      }
      q.child[k] = -1;
      q.count[k] = 0;
      continue;
    }
    const Node& c = nodes[kids[k]];
    for (int a = 0; a < 3; a++) {
      q.lo[a][k] = c.b.lo[a];
      q.hi[a][k] = c.b.hi[a];
    }
    if (c.count) {
      q.child[k] = (int32_t)c.start;
      q.count[k] = c.count;
    } else {
      q.child[k] = Collapse(kids[k]);
      q.count[k] = 0;
    }
  }
  nodes4[idx] = q;
  return idx;
}

bool Accel::IntersectEntry(uint32_t i, const Ray& ray, Float tnear, Float tfar, HitShapeRecord& rec) const {
  if (!isTri[i]) return (*shapes)[prims[i].shape]->IntersectPrim(prims[i].prim, ray, tnear, tfar, rec);
  const float* v = &triVerts[9 * size_t(i)];
  const V3 p0(v[0], v[1], v[2]), p1(v[3], v[4], v[5]), p2(v[6], v[7], v[8]);
  Float t, u, w;
  if (!IntersectTriangleMT(p0, p1, p2, ray, tnear, tfar, t, u, w)) return false;
  rec.T = t;
  rec.PrimitiveUV = V2(u, w);
  rec.PrimitiveIndex = prims[i].prim;
  rec.ShapeIndex = prims[i].shape;
  rec.GeometryNormal = Cross(p1 - p0, p2 - p0);
  return true;
}

namespace {
// Entry and exit distances of a ray against the four child boxes of a node.
// A lane whose slab product is 0 * inf (a ray parallel to a face plane that
// holds its origin) keeps that axis open, as the scalar slab test does.
struct RaySimd {
  __m128 o[3], inv[3], tmin;
};

inline RaySimd LoadRay(const Ray& ray) {
  RaySimd r;
  const float inv[3] = {1 / ray.D.x, 1 / ray.D.y, 1 / ray.D.z};
  for (int a = 0; a < 3; a++) {
    r.o[a] = _mm_set1_ps(ray.O[a]);
    r.inv[a] = _mm_set1_ps(inv[a]);
  }
  r.tmin = _mm_set1_ps(ray.MinT);
  return r;
}

inline int SlabTest4(const Accel::Node4& n, const RaySimd& r, float tfar, __m128& tnear) {
  const __m128 ninf = _mm_set1_ps(-Inf), pinf = _mm_set1_ps(Inf);
  __m128 tn = r.tmin, tf = pinf;
  for (int a = 0; a < 3; a++) {
    const __m128 t0 = _mm_mul_ps(_mm_sub_ps(_mm_load_ps(n.lo[a]), r.o[a]), r.inv[a]);
    const __m128 t1 = _mm_mul_ps(_mm_sub_ps(_mm_load_ps(n.hi[a]), r.o[a]), r.inv[a]);
    const __m128 nan = _mm_cmpunord_ps(t0, t1);
    tn = _mm_max_ps(tn, _mm_blendv_ps(_mm_min_ps(t0, t1), ninf, nan));
    tf = _mm_min_ps(tf, _mm_blendv_ps(_mm_max_ps(t0, t1), pinf, nan));
  }
  // robust slab test (Ize 2013): widen the exit by 3 * 2 ulp
  tf = _mm_min_ps(_mm_mul_ps(tf, _mm_set1_ps(1 + 2 * 1.2e-7f * 3)), _mm_set1_ps(tfar));
  tnear = tn;
  return _mm_movemask_ps(_mm_cmple_ps(tn, tf));
}
}  // namespace

bool Accel::Closest(const Ray& ray, HitShapeRecord& rec) const {
  if (nodes4.empty()) return false;
  const RaySimd r = LoadRay(ray);
  Float tfar = ray.MaxT;
  bool hit = false;
  struct Entry {
    int32_t node;
    float t;
  } stack[256];
  int sp = 0;
  stack[sp++] = {0, ray.MinT};
  while (sp) {
    const Entry e = stack[--sp];
    if (e.t > tfar) continue;
    const Node4& n = nodes4[e.node];
    __m128 tn4;
    int mask = SlabTest4(n, r, tfar, tn4);
    alignas(16) float tn[4];
    _mm_store_ps(tn, tn4);
    // leaves first: they can only shorten tfar
    for (int k = 0; k < 4; k++) {
      if (!(mask >> k & 1) || n.child[k] < 0 || n.count[k] == 0) continue;
      for (uint32_t i = uint32_t(n.child[k]), end = i + n.count[k]; i < end; i++) {
        HitShapeRecord h;
        if (IntersectEntry(i, ray, ray.MinT, tfar, h)) {
          rec = h;
          tfar = h.T;
          hit = true;
        }
      }
    }
    // inner children still in front of the closest hit, nearest popped first
    Entry in[4];
    int m = 0;
    for (int k = 0; k < 4; k++) {
      if (!(mask >> k & 1) || n.child[k] < 0 || n.count[k] != 0 || tn[k] > tfar) continue;
      Entry x{n.child[k], tn[k]};
      int j = m++;
      while (j > 0 && in[j - 1].t < x.t) {
        in[j] = in[j - 1];
        j--;
      }
      in[j] = x;  // descending distance
    }
    for (int k = 0; k < m; k++) stack[sp++] = in[k];
  }
  return hit;
}

bool Accel::Occluded(const Ray& ray) const {
  if (nodes4.empty()) return false;
  const RaySimd r = LoadRay(ray);
  int32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const Node4& n = nodes4[stack[--sp]];
    __m128 tn4;
    const int mask = SlabTest4(n, r, ray.MaxT, tn4);
    for (int k = 0; k < 4; k++) {
      if (!(mask >> k & 1) || n.child[k] < 0) continue;
      if (n.count[k] == 0) {
        stack[sp++] = n.child[k];
        continue;
      }
      for (uint32_t i = uint32_t(n.child[k]), end = i + n.count[k]; i < end; i++) {
        HitShapeRecord h;
        if (IntersectEntry(i, ray, ray.MinT, ray.MaxT, h)) return true;
      }
    }
  }
  return false;
}

bool Accel::ClosestBrute(const Ray& ray, HitShapeRecord& rec) const {
  Float tfar = ray.MaxT;
  bool hit = false;
  for (const auto& p : prims) {
    HitShapeRecord r;
    if ((*shapes)[p.shape]->IntersectPrim(p.prim, ray, ray.MinT, tfar, r)) {
      rec = r;
      tfar = r.T;
      hit = true;
    }
  }
  return hit;
}

bool Accel::OccludedBrute(const Ray& ray) const {
  for (const auto& p : prims) {
    HitShapeRecord r;
    if ((*shapes)[p.shape]->IntersectPrim(p.prim, ray, ray.MinT, ray.MaxT, r)) return true;
  }
  return false;
}

// Conservative segment-box test in double precision: the box is padded by
// 1e-4 of its extent plus magnitude and the segment by 1e-5 relative, far
// beyond the float rounding of any primitive test, so a primitive that reports
// a hit in [tmin, tmax] never sits in a rejected chunk.
static bool SegmentMayHit(const Bounds& b, const Ray& ray, Float tmax) {
  double t0 = double(ray.MinT) - 1e-5 * (1.0 + std::fabs(double(ray.MinT)));
  double t1 = double(tmax) + 1e-5 * (1.0 + std::fabs(double(tmax)));
  for (int a = 0; a < 3; a++) {
    double lo = b.lo[a], hi = b.hi[a];
    double pad = 1e-4 * ((hi - lo) + std::fabs(lo) + std::fabs(hi) + 1.0);
    lo -= pad;
    hi += pad;
    double o = ray.O[a], d = ray.D[a];
    if (d == 0.0) {
      if (o < lo || o > hi) return false;
      continue;
    }
    double tn = (lo - o) / d, tf = (hi - o) / d;
    if (tn > tf) std::swap(tn, tf);
    t0 = std::max(t0, tn);
    t1 = std::min(t1, tf);
    if (!(t0 <= t1)) return false;
  }
  return true;
}

bool Accel::ClosestCulled(const Ray& ray, HitShapeRecord& rec) const {
  Float tfar = ray.MaxT;
  bool hit = false;
  for (uint32_t c = 0; c < chunks.size(); c++) {
    if (!SegmentMayHit(chunks[c], ray, tfar)) continue;
    const uint32_t end = std::min<uint32_t>((c + 1) * kCullChunk, (uint32_t)prims.size());
    for (uint32_t i = c * kCullChunk; i < end; i++) {
      HitShapeRecord r;
      if ((*shapes)[prims[i].shape]->IntersectPrim(prims[i].prim, ray, ray.MinT, tfar, r)) {
        rec = r;
        tfar = r.T;
        hit = true;
      }
    }
  }
  return hit;
}

bool Accel::OccludedCulled(const Ray& ray) const {
  for (uint32_t c = 0; c < chunks.size(); c++) {
    if (!SegmentMayHit(chunks[c], ray, ray.MaxT)) continue;
    const uint32_t end = std::min<uint32_t>((c + 1) * kCullChunk, (uint32_t)prims.size());
    for (uint32_t i = c * kCullChunk; i < end; i++) {
      HitShapeRecord r;
      if ((*shapes)[prims[i].shape]->IntersectPrim(prims[i].prim, ray, ray.MinT, ray.MaxT, r)) return true;
    }
  }
  return false;
}

}  // namespace ref
