// ORACLE / TEST INFRASTRUCTURE ONLY (see ref_math.h).
// Scene objects of the restatement: interactions, shapes, BSDFs, lights,
// media, camera, sampler, accel.  Each piece cites the reference code it
// restates; paths are relative to module/rad.offline/ unless noted.
#pragma once

#include <memory>
#include <random>
#include <vector>

#include "../../include/rad_hip.h"
#include "ref_math.h"

namespace ref {

constexpr Float Inf = std::numeric_limits<Float>::infinity();

struct Shape;
struct Medium;
struct Bsdf;
struct Light;

// interaction.h:15-45, interaction.cpp:10-30
struct Interaction {
  Float T = Inf;
  V3 P, N;
  V3 OffsetP(const V3& d) const {
    Float mag = (1 + MaxCoeff(Abs(P))) * RayEpsilon;
    mag = MulSign(mag, Dot(N, d));
    return Fmadd(V3(mag), N, P);
  }
  Ray SpawnRay(const V3& d) const { return Ray{OffsetP(d), d, 0, std::numeric_limits<Float>::max()}; }
  bool IsValid() const { return T != Inf; }
};

struct PositionSampleResult {
  V3 P, N;
  V2 UV;
  Float Pdf = 0;
  bool IsDelta = false;
};
struct DirectionSampleResult : PositionSampleResult {
  V3 Dir;
  Float Dist = 0;
};

// interaction.h:47-100
struct SurfaceInteraction : Interaction {
  const Shape* shape = nullptr;
  uint32_t ShapeIndex = 0xffffffffu;
  Frame Shading;
  V2 UV;
  V3 dPdU, dPdV, dNdU, dNdV;
  V2 dUVdX, dUVdY;  // texture-space footprint (camera rays only, interaction.cpp:83-111)
  V3 Wi;
  void ComputeShadingFrame() {  // interaction.cpp:45-53
    if (IsZero(dPdU)) {
      Shading.S = CoordinateSystem(Shading.N).first;
    } else {
      Shading.S = Normalized(Shading.N * (-Dot(Shading.N, dPdU)) + dPdU);
    }
    Shading.T = Cross(Shading.N, Shading.S);
  }
  V3 ToWorld(const V3& v) const { return Shading.ToWorld(v); }
  V3 ToLocal(const V3& v) const { return Shading.ToLocal(v); }
  DirectionSampleResult ToDsr(const Interaction& ref) const {  // interaction.cpp:67-82
    DirectionSampleResult dsr{};
    dsr.P = P;
    dsr.N = Shading.N;
    dsr.UV = UV;
    dsr.Pdf = 0;
    dsr.IsDelta = false;
    V3 rel = P - ref.P;
    dsr.Dist = Norm(rel);
    dsr.Dir = IsValid() ? rel / dsr.Dist : -Wi;
    return dsr;
  }
  const Medium* GetMedium(const Ray& ray) const;
};

struct MediumInteraction : Interaction {  // interaction.h:102-111
  const Medium* medium = nullptr;
  Frame Shading;
  V3 Wi;
  Float MinT = 0;
  Spectrum SigmaS, SigmaN, SigmaT, CombinedExtinction;
};

struct HitShapeRecord {  // interaction.h:113-122
  Float T = Inf;
  V3 GeometryNormal;
  V2 PrimitiveUV;
  uint32_t PrimitiveIndex = 0;
  uint32_t ShapeIndex = 0xffffffffu;
};

// warp.cpp:10-99
namespace Warp {
inline V2 SquareToUniformDisk(const V2& u) {
  Float r = std::sqrt(u.x);
  Float angle = 2 * PI * u.y;
  return V2(r * std::cos(angle), r * std::sin(angle));
}
inline V3 SquareToCosineHemisphere(const V2& u) {
  V2 b = SquareToUniformDisk(u);
  Float z = SafeSqrt(1 - Sqr(b.x) - Sqr(b.y));
  return V3(b.x, b.y, z);
}
inline Float SquareToCosineHemispherePdf(const V3& v) { return v.z * (1 / PI); }
inline V2 SquareToUniformTriangle(const V2& u) {
  Float t = SafeSqrt(1 - u.x);
  return V2(1 - t, t * u.y);
}
inline V3 SquareToUniformSphere(const V2& u) {
  Float z = Fmadd(-2, u.y, 1);
  Float r = SafeSqrt(Fmadd(-z, z, 1));
  Float a = 2 * PI * u.x;
  return V3(r * std::cos(a), r * std::sin(a), z);
}
inline Float SquareToUniformSpherePdf() { return 1 / (4 * PI); }
inline Float SquareToUniformConePdf(Float cosCutoff) { return 1 / (2 * PI) / (1 - cosCutoff); }
}  // namespace Warp

// distribution.cpp:7-53
struct DiscreteDistribution1D {
  std::vector<Float> pmf, cdf;
  Float sum = 0, normalization = 0;
  void Build(const std::vector<Float>& data) {
    pmf = data;
    cdf.resize(data.size());
    double s = 0;
    for (size_t i = 0; i < data.size(); i++) {
      s += (double)data[i];
      cdf[i] = (Float)s;
    }
    sum = Float(s);
    normalization = Float(1.0 / s);
  }
  size_t Sample(Float xi) const {
    Float value = xi * sum;
    return size_t(std::lower_bound(cdf.begin(), cdf.end(), value) - cdf.begin());
  }
  std::pair<size_t, Float> SampleReuse(Float xi) const {
    size_t index = Sample(xi);
    Float p = pmf[index] * normalization;
    Float c = index == 0 ? 0 : cdf[index - 1] * normalization;
    return {index, (xi - c) / p};
  }
};

// distribution.cpp:55-128 (ContinuousDistribution1D / 2D)
struct Dist1D {
  std::vector<Float> pdf, cdf;
  Float sum = 0, normalization = 0;
  void Build(const Float* p, size_t n);
  size_t Count() const { return pdf.size(); }
  // (cell index, continuous position, normalised density)
  std::tuple<size_t, Float, Float> Sample(Float xi) const;
};
struct Dist2D {
  std::vector<Dist1D> conditional;
  Dist1D marginal;
  void Build(const Float* data, size_t nu, size_t nv);
  // (cell u, cell v, density over [0,1]^2, continuous position in [0,1]^2)
  std::tuple<size_t, size_t, Float, V2> Sample(const V2& xi) const;
  Float Pdf(const V2& uv) const;
};

struct Bounds {
  V3 lo{Inf, Inf, Inf}, hi{-Inf, -Inf, -Inf};
  void Extend(const V3& p) { lo = Min(lo, p); hi = Max(hi, p); }
  void Extend(const Bounds& b) { lo = Min(lo, b.lo); hi = Max(hi, b.hi); }
  V3 Center() const { return (lo + hi) * 0.5f; }
  Float Area() const {
    V3 d = hi - lo;
    if (d.x < 0) return 0;
    return 2 * (d.x * d.y + d.y * d.z + d.z * d.x);
  }
};

// ---- shapes (shape.h:15-101) ----
struct Shape {
  Float surfaceArea = 0;
  const Bsdf* bsdf = nullptr;
  int32_t bsdfIndex = -1;
  const Light* light = nullptr;
  const Medium* inside = nullptr;
  const Medium* outside = nullptr;
  uint32_t index = 0;
  virtual ~Shape() = default;
  virtual uint32_t PrimCount() const = 0;
  virtual Bounds PrimBounds(uint32_t prim) const = 0;
  // Candidate hit against the current [near, far]; returns true and the hit.
  virtual bool IntersectPrim(uint32_t prim, const Ray& ray, Float near, Float far, HitShapeRecord& rec) const = 0;
  virtual SurfaceInteraction ComputeInteraction(const Ray& ray, const HitShapeRecord& rec) const = 0;
  virtual PositionSampleResult SamplePosition(const V2& xi) const = 0;
  virtual Float PdfPosition(const PositionSampleResult& psr) const = 0;
  virtual DirectionSampleResult SampleDirection(const Interaction& ref, const V2& xi) const;
  virtual Float PdfDirection(const Interaction& ref, const DirectionSampleResult& dsr) const;
  bool IsLight() const { return light != nullptr; }
  bool HasBsdf() const { return bsdf != nullptr; }
  bool HasMedia() const { return inside != nullptr || outside != nullptr; }
};

// Triangle test shared bit-for-bit with the device kernels: the Moller-
// Trumbore variant Embree 3 applies to its TriangleM primitives
// (e1 = v0 - v1, e2 = v2 - v0, Ng = e2 x e1; accept den != 0, U, V >= 0,
// U + V <= |den|, |den|*tnear < T <= |den|*tfar).  No FMA contraction.
bool IntersectTriangleMT(const V3& v0, const V3& v1, const V3& v2, const Ray& ray, Float tnear, Float tfar,
                         Float& t, Float& u, Float& v);

struct MeshShape : Shape {  // shape/mesh_base.cpp, shape/mesh.cpp, shape/cube.cpp
  Transform toWorld;
  std::vector<V3> position, normal;
  std::vector<V2> uv;
  std::vector<uint32_t> indices;
  uint32_t triangleCount = 0;
  DiscreteDistribution1D dist;
  uint32_t PrimCount() const override { return triangleCount; }
  Bounds PrimBounds(uint32_t prim) const override;
  bool IntersectPrim(uint32_t prim, const Ray& ray, Float near, Float far, HitShapeRecord& rec) const override;
  SurfaceInteraction ComputeInteraction(const Ray& ray, const HitShapeRecord& rec) const override;
  PositionSampleResult SamplePosition(const V2& xi) const override;
  Float PdfPosition(const PositionSampleResult&) const override { return dist.normalization; }
  void UpdateDistribution();
};

struct SphereShape : Shape {  // shape/sphere.cpp
  V3 center;
  Float radius = 1;
  Transform toWorld;
  uint32_t PrimCount() const override { return 1; }
  Bounds PrimBounds(uint32_t) const override;
  bool IntersectPrim(uint32_t prim, const Ray& ray, Float near, Float far, HitShapeRecord& rec) const override;
  SurfaceInteraction ComputeInteraction(const Ray& ray, const HitShapeRecord& rec) const override;
  PositionSampleResult SamplePosition(const V2& xi) const override;
  Float PdfPosition(const PositionSampleResult&) const override { return 1 / surfaceArea; }
  DirectionSampleResult SampleDirection(const Interaction& ref, const V2& xi) const override;
  Float PdfDirection(const Interaction& ref, const DirectionSampleResult& dsr) const override;
};

struct RectangleShape : Shape {  // shape/rectangle.cpp
  Transform toWorld;
  Frame frame;
  float toObject[4][4];
  uint32_t PrimCount() const override { return 1; }
  Bounds PrimBounds(uint32_t) const override;
  bool IntersectPrim(uint32_t prim, const Ray& ray, Float near, Float far, HitShapeRecord& rec) const override;
  SurfaceInteraction ComputeInteraction(const Ray& ray, const HitShapeRecord& rec) const override;
  PositionSampleResult SamplePosition(const V2& xi) const override;
  Float PdfPosition(const PositionSampleResult&) const override { return 1 / surfaceArea; }
};

// ---- BSDF (render/bsdf.h) ----
enum BsdfType : uint32_t {
  BSDF_DIFFUSE = 1, BSDF_GLOSSY = 2, BSDF_DELTA = 4, BSDF_REFLECTION = 8, BSDF_TRANSMISSION = 16,
  BSDF_NO_DELTA = BSDF_DIFFUSE | BSDF_GLOSSY
};
struct BsdfSampleResult {
  V3 Wo;
  Float Pdf = 0;
  Float Eta = 1;
  uint32_t TypeMask = 0;
};
struct Bsdf {
  uint32_t flags = 0;
  bool quirks = true;
  virtual ~Bsdf() = default;
  bool HasAnyTypeExceptDelta() const { return (flags & BSDF_NO_DELTA) != 0; }
  virtual std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float lobeXi, const V2& dirXi) const = 0;
  virtual Spectrum Eval(const SurfaceInteraction& si, const V3& wo) const = 0;
  virtual Float Pdf(const SurfaceInteraction& si, const V3& wo) const = 0;
};
std::unique_ptr<Bsdf> MakeBsdf(const rad_bsdf_desc& d, bool quirks);
// bsdfs[index] of a scene's array (two_side / mask resolve their nested entry)
// specWeights (optional, per array entry, NaN = none): the plastic specular
// sampling weight taken from textured diffuse / specular parameters.
std::unique_ptr<Bsdf> MakeBsdf(const rad_bsdf_desc* all, uint32_t count, uint32_t index, bool quirks, int depth = 0,
                               const Float* specWeights = nullptr);
float RefErfInv(float a);

// ---- textures (render/texture.h, texture/bitmap.cpp, texture/chessboard.cpp) ----
struct TexImage {  // BlockArray2D<T> (the block tiling is only a storage layout); mono in .x
  uint32_t w = 0, h = 0;
  std::vector<V3> px;
  const V3& At(uint32_t x, uint32_t y) const;
};
struct Texture {
  int32_t type = RAD_TEXTURE_BITMAP, channels = 3;
  // bitmap
  std::vector<TexImage> mips;
  int32_t filter = RAD_FILTER_BILINEAR, wrap = RAD_WRAP_CLAMP;
  Float aniso = 0;
  // chessboard
  Transform toUV;
  const Texture* a = nullptr;
  const Texture* b = nullptr;
  V3 colorA, colorB;
  bool remap = false;

  V3 Eval(const V2& uv, const V2& dx, const V2& dy) const;
  double MeanLuminance() const;
  void BuildMips(TexImage level0, int32_t maxLevel);
  V3 BilinearAt(const TexImage& img, uint32_t u, uint32_t v) const;
  V3 Bilinear4(const TexImage& img, uint32_t pu, uint32_t pv, Float fu, Float fv) const;
  V3 Nearest(const TexImage& img, const V2& uv) const;
  V3 Bilinear(const TexImage& img, const V2& uv) const;
  V3 Isotropic(const V2& uv, size_t level) const;
  V3 Ewa(const V2& uv, size_t level, const V2& d0, const V2& d1) const;
  V3 Trilinear(const V2& uv, const V2& dx, const V2& dy) const;
  V3 Anisotropic(const V2& uv, const V2& dx, const V2& dy) const;
};
void BuildTextures(const rad_texture_desc* descs, uint32_t count, std::vector<std::unique_ptr<Texture>>& out);

// A rectangle's parametric surface (rectangle.cpp:236-250).
SurfaceInteraction RectangleParamSurface(const Shape* rect, const V2& uv);

// RayDifferential of a camera ray (ray.h:35-58, perspective.cpp:84-99).
struct RayDiff {
  bool has = false;
  V3 Ox, Oy, Dx, Dy;
};
void ComputeUVDifferentials(SurfaceInteraction& si, const RayDiff& rd);

// ---- lights (render/light.h) ----
struct Scene;
struct Light {
  const Shape* shape = nullptr;
  virtual ~Light() = default;
  virtual bool IsEnv() const { return false; }
  virtual Spectrum Eval(const SurfaceInteraction& si) const = 0;
  virtual std::pair<DirectionSampleResult, Spectrum> SampleDirection(const Interaction& ref, const V2& xi) const = 0;
  virtual Float PdfDirection(const Interaction& ref, const DirectionSampleResult& dsr) const = 0;
  virtual void SetScene(const Scene&) {}
};

// ---- media (render/medium.h, medium.cpp) ----
struct Medium {
  int phaseType = RAD_PHASE_ISOTROPIC;
  Float g = 0.8f;
  bool isHomogeneous = true;
  virtual ~Medium() = default;
  virtual std::tuple<bool, Float, Float> IntersectAABB(const Ray& ray) const = 0;
  virtual Spectrum GetMajorant(const MediumInteraction& mi) const = 0;
  virtual std::tuple<Spectrum, Spectrum, Spectrum> GetScatteringCoefficients(const MediumInteraction& mi) const = 0;
  MediumInteraction SampleInteraction(const Ray& ray, Float sample, uint32_t channel) const;
  std::pair<Spectrum, Spectrum> EvalTrAndPdf(const MediumInteraction& mi, const SurfaceInteraction& si) const;
  // phase_function/henyey_greenstein.cpp:18-40, isotropic.cpp:16-24
  std::pair<V3, Float> PhaseSample(const MediumInteraction& mi, const V2& xi) const;
  Float PhaseEval(const MediumInteraction& mi, const V3& wo) const;
};

// volume/grid.cpp:41-103 (sigma_t volume of a heterogeneous medium)
struct GridVolume {
  int sx = 0, sy = 0, sz = 0, channels = 1, wrap = RAD_WRAP_CLAMP;
  std::vector<float> data;
  Float maxValue = 0;
  Spectrum Read(int u, int v, int w) const;
  Spectrum Eval(const V3& p) const;
};

std::unique_ptr<Medium> MakeMedium(const rad_medium_desc& d);

// ---- camera (camera/perspective.cpp) ----
struct Camera {
  Float nearClip = 0.001f, farClip = 100;
  int width = 1280, height = 720;
  V2 rcpResolution;
  Transform cameraToWorld, cameraToClip;
  V3 dx, dy;  // camera-space offset of one pixel on the near plane
  void Init(const rad_camera_desc& d);
  Ray SampleRay(const V2& screenPosition) const;
  Ray SampleRayDifferential(const V2& screenPosition, RayDiff& rd) const;
};

// ---- sampler (sampler/independent.cpp) ----
struct Sampler {
  std::mt19937 engine;
  std::uniform_real_distribution<Float> dist;
  explicit Sampler(uint32_t seed) : engine(seed) {}
  Float Next1D() { return dist(engine); }
  V2 Next2D() {
    Float a = dist(engine);
    Float b = dist(engine);
    return V2(a, b);
  }
};

// ---- accel: binned-SAH BVH over every primitive of every shape ----
struct PrimRef {
  uint32_t shape, prim;
};
// CPU accel: a binned-SAH binary tree collapsed into a 4-wide tree whose
// nodes hold their children's boxes as SoA lanes, so one SSE slab test
// covers a node's four children (the timed CPU path; Embree itself is not
// available).  Leaf triangles are tested inline from a packed copy of their
// vertices with the same arithmetic as MeshShape::IntersectPrim.
struct Accel {
  struct Node {
    Bounds b;
    int32_t left = -1, right = -1;
    uint32_t start = 0, count = 0;
  };
  struct alignas(64) Node4 {
    float lo[3][4], hi[3][4];  // child boxes, axis-major, one lane per child
    int32_t child[4];          // inner: node index; leaf: first entry of `prims`; empty: -1
    uint32_t count[4];         // 0: inner; > 0: leaf primitive count
  };
  std::vector<Node> nodes;       // binary build tree
  std::vector<Node4> nodes4;     // traversed tree
  std::vector<float> triVerts;   // 9 floats per entry of `prims` (mesh triangles)
  std::vector<uint8_t> isTri;    // entry of `prims` is a mesh triangle
  std::vector<PrimRef> prims;
  const std::vector<std::unique_ptr<Shape>>* shapes = nullptr;
  Bounds world;
  double buildMs = 0;
  void Build(const std::vector<std::unique_ptr<Shape>>& shapes);
  bool Closest(const Ray& ray, HitShapeRecord& rec) const;
  bool Occluded(const Ray& ray) const;
  bool ClosestBrute(const Ray& ray, HitShapeRecord& rec) const;
  bool OccludedBrute(const Ray& ray) const;
  // Brute force over every primitive in the order of `prims`, skipping runs of
  // kCullChunk consecutive primitives whose padded box the ray segment misses
  // (a primitive hit at t lies inside its own box, so the result equals
  // ClosestBrute / OccludedBrute; tests check that bit for bit).
  static constexpr uint32_t kCullChunk = 64;
  std::vector<Bounds> chunks;
  bool ClosestCulled(const Ray& ray, HitShapeRecord& rec) const;
  bool OccludedCulled(const Ray& ray) const;
  int BuildRecursive(std::vector<PrimRef>& refs, std::vector<Bounds>& pb, std::vector<V3>& pc, uint32_t begin, uint32_t end, int depth);
  int Collapse(int binaryNode);
  bool IntersectEntry(uint32_t i, const Ray& ray, Float tnear, Float tfar, HitShapeRecord& rec) const;
};

// ---- scene (scene.cpp:11-122) ----
struct Scene {
  std::vector<std::unique_ptr<Shape>> shapes;
  std::vector<std::unique_ptr<Bsdf>> bsdfs;
  std::vector<std::unique_ptr<Light>> lights;
  std::vector<std::unique_ptr<Medium>> media;
  const Medium* globalMedium = nullptr;
  const Light* envLight = nullptr;
  Float lightPdf = 0;
  Camera camera;
  Accel accel;
  bool quirks = true;
  bool brute = false;
  // textured parameters: the BSDF descriptors, which of them (or their nested
  // BSDFs) read a texture, and the plastic weights from texel means
  std::vector<rad_bsdf_desc> bsdfDescs;
  std::vector<bool> bsdfTextured;
  std::vector<Float> specWeights;
  std::vector<std::unique_ptr<Texture>> textures;

  void Create(const rad_scene_desc& desc, bool quirks);
  const Bsdf* HitBsdf(const SurfaceInteraction& si, std::unique_ptr<Bsdf>& holder) const;
  bool RayIntersect(const Ray& ray) const { return brute ? accel.OccludedBrute(ray) : accel.Occluded(ray); }
  bool RayIntersect(const Ray& ray, SurfaceInteraction& si) const;
  bool IsOcclude(const Interaction& ref, const V3& p) const;
  std::pair<uint32_t, Float> SampleLight(Float xi) const;
  std::tuple<const Light*, DirectionSampleResult, Spectrum> SampleLightDirection(const Interaction& ref, Float sl, const V2& xi) const;
  Float PdfLightDirection(const Light* light, const Interaction& ref, const DirectionSampleResult& dsr) const {
    return light->PdfDirection(ref, dsr) * lightPdf;
  }
  const Light* GetLight(const SurfaceInteraction& si) const {
    return si.shape == nullptr ? envLight : (si.shape->IsLight() ? si.shape->light : nullptr);
  }
};

// ---- integrators (renderer/path.cpp, renderer/vol_path.cpp) ----
Spectrum PathLi(const Ray& ray, const Scene& scene, Sampler& sampler, int maxDepth, int rrDepth,
                const RayDiff& diff = RayDiff());
Spectrum VolPathLi(const Ray& ray, const Scene& scene, Sampler& sampler, int maxDepth, int rrDepth);

}  // namespace ref
