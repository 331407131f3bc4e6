// ORACLE / TEST INFRASTRUCTURE ONLY (see ref_math.h).
// Media, the `path` and `vol_path` integrators, the tile-parallel sample
// driver, and the extern "C" surface the tests and the CPU baseline call.
#include <omp.h>

#include <atomic>
#include <chrono>
#include <cstring>
#include <string>

#include "ref_scene.h"
#include "refcpu.h"

namespace ref {

// ---------------------------------------------------------------- grid volume (volume/grid.cpp:41-103)
Spectrum GridVolume::Read(int u, int v, int w) const {
  size_t start = (((size_t)w * (size_t)sy + (size_t)v) * (size_t)sx + (size_t)u) * (size_t)channels;
  if (channels == 1) return Spectrum(data[start]);
  return Spectrum(data[start], data[start + 1], data[start + 2]);
}

static Float WrapCoord(Float v, int wrap) {
  if (wrap == RAD_WRAP_REPEAT) return std::min(std::max(v - std::floor(v), Float(0)), Float(1));
  return std::min(std::max(v, Float(0)), Float(1));
}

Spectrum GridVolume::Eval(const V3& p) const {
  V3 wp(WrapCoord(p.x, wrap), WrapCoord(p.y, wrap), WrapCoord(p.z, wrap));
  V3 fp(wp.x * Float(sx - 1), wp.y * Float(sy - 1), wp.z * Float(sz - 1));
  int ipx = int(fp.x), ipy = int(fp.y), ipz = int(fp.z);
  int dpu = (fp.x > (ipx + Float(0.5))) ? 1 : -1;
  int dpv = (fp.y > (ipy + Float(0.5))) ? 1 : -1;
  int dpw = (fp.z > (ipz + Float(0.5))) ? 1 : -1;
  int apu = std::min(std::max(ipx + dpu, 0), sx - 1);
  int apv = std::min(std::max(ipy + dpv, 0), sy - 1);
  int apw = std::min(std::max(ipz + dpw, 0), sz - 1);
  Float du = std::min(std::abs(fp.x - ipx - Float(0.5)), Float(1));
  Float dv = std::min(std::abs(fp.y - ipy - Float(0.5)), Float(1));
  Float dw = std::min(std::abs(fp.z - ipz - Float(0.5)), Float(1));
  Spectrum u0v0w0 = Read(ipx, ipy, ipz), u1v0w0 = Read(apu, ipy, ipz), u0v1w0 = Read(ipx, apv, ipz),
           u0v0w1 = Read(ipx, ipy, apw), u1v1w0 = Read(apu, apv, ipz), u0v1w1 = Read(ipx, apv, apw),
           u1v0w1 = Read(apu, ipy, apw), u1v1w1 = Read(apu, apv, apw);
  V3 du_v0w0 = Fmadd(u0v0w0, V3(1 - du), u1v0w0 * du);
  V3 du_v1w0 = Fmadd(u0v1w0, V3(1 - du), u1v1w0 * du);
  V3 dv_w0 = Fmadd(du_v0w0, V3(1 - dv), du_v1w0 * dv);
  V3 du_v0w1 = Fmadd(u0v0w1, V3(1 - du), u1v0w1 * du);
  V3 du_v1w1 = Fmadd(u0v1w1, V3(1 - du), u1v1w1 * du);
  V3 dv_w1 = Fmadd(du_v0w1, V3(1 - dv), du_v1w1 * dv);
  return Fmadd(dv_w0, V3(1 - dw), dv_w1 * dw);
}

// ---------------------------------------------------------------- media
struct HomogeneousMedium final : Medium {  // medium/homogeneous.cpp:14-41
  Spectrum sigmaT, albedo;
  Float scale = 1;
  std::tuple<bool, Float, Float> IntersectAABB(const Ray&) const override { return {true, 0.f, Inf}; }
  Spectrum EvalSigmaT() const { return sigmaT * scale; }
  Spectrum GetMajorant(const MediumInteraction&) const override { return EvalSigmaT(); }
  std::tuple<Spectrum, Spectrum, Spectrum> GetScatteringCoefficients(const MediumInteraction&) const override {
    Spectrum st = EvalSigmaT();
    return {Mul(st, albedo), Spectrum(0), st};
  }
};

struct HeterogeneousMedium final : Medium {  // medium/heterogeneous.cpp:15-79
  GridVolume grid;
  bool hasGrid = false;
  Spectrum constSigmaT, albedo;
  Float scale = 1, maxDensity = 0;
  Transform toWorld;
  Bounds bbox;
  std::tuple<bool, Float, Float> IntersectAABB(const Ray& ray) const override {
    bool active = !IsZero(ray.D) ||
                  (ray.O.x >= bbox.lo.x && ray.O.y >= bbox.lo.y && ray.O.z >= bbox.lo.z && ray.O.x <= bbox.hi.x &&
                   ray.O.y <= bbox.hi.y && ray.O.z <= bbox.hi.z);
    V3 dRcp(1 / ray.D.x, 1 / ray.D.y, 1 / ray.D.z);
    V3 t1 = Mul(bbox.lo - ray.O, dRcp), t2 = Mul(bbox.hi - ray.O, dRcp);
    V3 t1p = Min(t1, t2), t2p = Max(t1, t2);
    Float mint = MaxCoeff(t1p), maxt = MinCoeff(t2p);
    active = active && maxt >= mint;
    return {active, mint, maxt};
  }
  Spectrum GetMajorant(const MediumInteraction&) const override { return Spectrum(maxDensity); }
  Spectrum EvalSigmaT(const MediumInteraction& mi) const {
    V3 p = toWorld.ApplyAffineToLocal(mi.P);
    Spectrum e = hasGrid ? grid.Eval(p) : constSigmaT;
    return e * scale;
  }
  std::tuple<Spectrum, Spectrum, Spectrum> GetScatteringCoefficients(const MediumInteraction& mi) const override {
    Spectrum st = EvalSigmaT(mi);
    return {Mul(st, albedo), Spectrum(maxDensity) - st, st};
  }
};

std::unique_ptr<Medium> MakeMedium(const rad_medium_desc& d) {
  std::unique_ptr<Medium> out;
  Spectrum albedo(d.albedo[0], d.albedo[1], d.albedo[2]);
  if (d.type == RAD_MEDIUM_HOMOGENEOUS) {
    auto m = std::make_unique<HomogeneousMedium>();
    m->sigmaT = Spectrum(d.sigma_t[0], d.sigma_t[1], d.sigma_t[2]);
    m->albedo = albedo;
    m->scale = d.scale;
    m->isHomogeneous = true;
    out = std::move(m);
  } else {
    auto m = std::make_unique<HeterogeneousMedium>();
    Float A[16] = {2, 0, 0, -1, 0, 2, 0, -1, 0, 0, 2, -1, 0, 0, 0, 1};  // S(2) * T(-0.5)
    m->toWorld = Transform(Mat4::FromRowMajor(d.to_world) * Mat4::FromRowMajor(A));
    for (int c = 0; c < 8; c++) {  // ApplyBoxToWorld of [0,1]^3 (transform.cpp:52-61)
      V3 p(c & 1 ? 1.f : 0.f, c & 2 ? 1.f : 0.f, c & 4 ? 1.f : 0.f);
      m->bbox.Extend(m->toWorld.ApplyAffineToWorld(p));
    }
    m->albedo = albedo;
    m->scale = d.scale;
    Float maxValue;
    if (d.grid && d.grid_size[0] > 0) {
      m->hasGrid = true;
      m->grid.sx = d.grid_size[0];
      m->grid.sy = d.grid_size[1];
      m->grid.sz = d.grid_size[2];
      m->grid.channels = d.grid_channels;
      m->grid.wrap = d.grid_wrap;
      size_t n = size_t(d.grid_size[0]) * d.grid_size[1] * d.grid_size[2] * d.grid_channels;
      m->grid.data.assign(d.grid, d.grid + n);
      maxValue = -Inf;
      for (float v : m->grid.data) maxValue = std::max(maxValue, v);
    } else {
      m->constSigmaT = Spectrum(d.sigma_t[0], d.sigma_t[1], d.sigma_t[2]);
      maxValue = MaxCoeff(m->constSigmaT);
    }
    m->maxDensity = maxValue * m->scale;
    m->isHomogeneous = false;
    out = std::move(m);
  }
  out->phaseType = d.phase;
  out->g = d.g;
  return out;
}

// medium.cpp:19-43
MediumInteraction Medium::SampleInteraction(const Ray& ray, Float sample, uint32_t channel) const {
  MediumInteraction mei{};
  mei.Wi = -ray.D;
  mei.Shading = Frame(mei.Wi);
  auto [aabbIts, mint, maxt] = IntersectAABB(ray);
  aabbIts &= (std::isfinite(mint) || std::isfinite(maxt));
  bool active = aabbIts;
  if (!active) {
    mint = 0;
    maxt = Inf;
  }
  mint = std::max(Float(0), mint);
  maxt = std::min(ray.MaxT, maxt);
  Spectrum combined = GetMajorant(mei);
  Float m = combined[channel];
  Float sampledT = mint + (-std::log(1 - sample) / m);
  bool validMi = active && (sampledT <= maxt);
  mei.T = validMi ? sampledT : Inf;
  mei.P = ray(sampledT);
  mei.medium = this;
  mei.MinT = mint;
  std::tie(mei.SigmaS, mei.SigmaN, mei.SigmaT) = GetScatteringCoefficients(mei);
  mei.CombinedExtinction = combined;
  return mei;
}

// medium.cpp:45-59
std::pair<Spectrum, Spectrum> Medium::EvalTrAndPdf(const MediumInteraction& mi, const SurfaceInteraction& si) const {
  Float t = std::min(mi.T, si.T) - mi.MinT;
  if (t <= 0) return {Spectrum(1), Spectrum(1)};
  Spectrum tr = Exp(mi.CombinedExtinction * -t);
  Spectrum pdf = si.T < mi.T ? tr : Mul(tr, mi.CombinedExtinction);
  if (tr.x == pdf.x && tr.y == pdf.y && tr.z == pdf.z) return {Spectrum(1), Spectrum(1)};
  return {tr, pdf};
}

static Float EvalHG(Float g, Float cosTheta) {
  Float temp = Float(1) + g * g + Float(2) * g * cosTheta;
  return (1 / (4 * PI)) * (1 - g * g) / (temp * std::sqrt(temp));
}

std::pair<V3, Float> Medium::PhaseSample(const MediumInteraction& mi, const V2& xi) const {
  if (phaseType == RAD_PHASE_ISOTROPIC) return {Warp::SquareToUniformSphere(xi), Warp::SquareToUniformSpherePdf()};
  Float sqrTerm = (1 - g * g) / (1 - g + 2 * g * xi.x);
  Float cosTheta = (1 + g * g - sqrTerm * sqrTerm) / (2 * g);
  if (std::abs(g) < Float32Epsilon) cosTheta = 1 - 2 * xi.x;
  Float sinTheta = SafeSqrt(1 - cosTheta * cosTheta);
  Float a = 2 * PI * xi.y;
  V3 wo(sinTheta * std::cos(a), sinTheta * std::sin(a), -cosTheta);
  wo = mi.Shading.ToWorld(wo);
  return {wo, EvalHG(g, -cosTheta)};
}

Float Medium::PhaseEval(const MediumInteraction& mi, const V3& wo) const {
  if (phaseType == RAD_PHASE_ISOTROPIC) return Warp::SquareToUniformSpherePdf();
  return EvalHG(g, Dot(wo, mi.Wi));
}

// ---------------------------------------------------------------- path (renderer/path.cpp:131-236)
static Float MisWeight(Float a, Float b) {
  a *= a;
  b *= b;
  Float w = a / (a + b);
  return std::isfinite(w) ? w : 0;
}

Spectrum PathLi(const Ray& ray0, const Scene& scene, Sampler& sampler, int maxDepth, int rrDepth, const RayDiff& diff) {
  Ray ray = ray0;
  RayDiff rd = diff;  // only the camera ray carries differentials (SpawnRay drops them)
  Spectrum throughput(1), result(0);
  Float eta = 1;
  int depth = 0;
  bool isSpecularPath = true;
  SurfaceInteraction prevSi{};
  Float prevBsdfPdf = 1;
  for (;; depth++) {
    SurfaceInteraction si{};
    bool anyHit = scene.RayIntersect(ray, si);
    const Light* hitLight = scene.GetLight(si);
    if (hitLight) {
      Float weight;
      if (isSpecularPath) {
        weight = 1;
      } else {
        DirectionSampleResult dsr = si.ToDsr(prevSi);
        Float lightPdf = scene.PdfLightDirection(hitLight, prevSi, dsr);
        weight = MisWeight(prevBsdfPdf, lightPdf);
      }
      Spectrum li = hitLight->Eval(si);
      result += Mul(throughput, li) * weight;
    }
    if (!anyHit || (maxDepth >= 0 && depth + 1 >= maxDepth)) break;
    if (si.shape->IsLight()) break;
    ComputeUVDifferentials(si, rd);  // si.BSDF(ray) (path.cpp:172)
    std::unique_ptr<Bsdf> resolved;
    const Bsdf* bsdf = scene.HitBsdf(si, resolved);
    if (bsdf->HasAnyTypeExceptDelta()) {
      Float sl = sampler.Next1D();
      V2 xi = sampler.Next2D();
      auto [l, dsr, li] = scene.SampleLightDirection(si, sl, xi);
      if (dsr.Pdf > 0) {
        V3 wo = si.ToLocal(dsr.Dir);
        Spectrum f = bsdf->Eval(si, wo);
        Float bsdfPdf = bsdf->Pdf(si, wo);
        if (!scene.IsOcclude(si, dsr.P)) {
          Float weight = dsr.IsDelta ? 1 : MisWeight(dsr.Pdf, bsdfPdf);
          result += Mul(Mul(throughput, f), li) * weight / dsr.Pdf;
        }
      }
    }
    {
      Float lobe = sampler.Next1D();
      V2 dxi = sampler.Next2D();
      auto [bsr, f] = bsdf->Sample(si, lobe, dxi);
      if (bsr.Pdf <= 0) break;
      throughput = Mul(throughput, f) / bsr.Pdf;
      eta *= bsr.Eta;
      ray = si.SpawnRay(si.ToWorld(bsr.Wo));
      rd.has = false;
      prevBsdfPdf = bsr.Pdf;
      prevSi = si;
      isSpecularPath = (bsr.TypeMask & BSDF_DELTA) != 0;
    }
    if (depth >= rrDepth) {
      Float rr = std::min(MaxCoeff(throughput) * Sqr(eta), Float(0.95));
      if (sampler.Next1D() > rr) break;
      throughput *= Rcp(rr);
    }
  }
  return result;
}

// ---------------------------------------------------------------- vol_path (renderer/vol_path.cpp:27-349)
static std::pair<Spectrum, DirectionSampleResult> VolSampleLight(const Scene& scene, Sampler& sampler, const Medium* medium,
                                                                 uint32_t channel, const Interaction& refIts) {
  Float sl = sampler.Next1D();
  V2 xi = sampler.Next2D();
  auto [l, dsr, li] = scene.SampleLightDirection(refIts, sl, xi);
  if (dsr.Pdf <= 0) return {Spectrum(0), dsr};
  Ray ray = refIts.SpawnRay(dsr.Dir);
  Float totalDist = 0;
  Spectrum resultTr(1);
  bool active = true;
  bool needsIntersection = true;
  SurfaceInteraction si{};
  while (active) {
    Float remainingDist = dsr.Dist * (1 - ShadowEpsilon) - totalDist;
    ray.MaxT = remainingDist;
    active &= remainingDist > 0;
    if (!active) break;
    bool escapedMedium = false;
    bool activeMedium = active && medium != nullptr;
    bool activeSurface = active && !activeMedium;
    if (activeMedium) {
      MediumInteraction mei = medium->SampleInteraction(ray, sampler.Next1D(), channel);
      if (medium->isHomogeneous && mei.IsValid()) ray.MaxT = std::min(mei.T, remainingDist);
      bool intersect = needsIntersection && activeMedium;
      if (intersect) scene.RayIntersect(ray, si);
      if (si.T < mei.T) mei.T = Inf;
      bool isSpectral = activeMedium;
      bool notSpectral = false;
      if (isSpectral) {
        Float t = std::min(remainingDist, std::min(mei.T, si.T)) - mei.MinT;
        if (t >= 0) {
          Spectrum tr = Exp(mei.CombinedExtinction * -t);
          Spectrum freeFlightPdf = (si.T < mei.T || mei.T > remainingDist) ? tr : Mul(tr, mei.CombinedExtinction);
          Float trPdf = freeFlightPdf[channel];
          Spectrum mTr = trPdf > 0 ? tr / trPdf : Spectrum(0);
          resultTr = Mul(resultTr, mTr);
        }
      }
      if (activeMedium && (mei.T > remainingDist) && mei.IsValid()) totalDist = dsr.Dist;
      if (activeMedium && (mei.T > remainingDist)) mei.T = Inf;
      escapedMedium = activeMedium && !mei.IsValid();
      activeMedium &= mei.IsValid();
      isSpectral &= activeMedium;
      notSpectral &= activeMedium;
      if (activeMedium) totalDist += mei.T;
      if (activeMedium) {
        ray.O = mei.P;
        si.T = si.T - mei.T;
        if (isSpectral) resultTr = Mul(resultTr, mei.SigmaN);
        if (notSpectral) resultTr = Div(Mul(resultTr, mei.SigmaN), mei.CombinedExtinction);
      }
    }
    bool intersect = activeSurface && needsIntersection;
    if (intersect) scene.RayIntersect(ray, si);
    needsIntersection &= !intersect;
    activeSurface |= escapedMedium;
    if (activeSurface) totalDist += si.T;
    activeSurface &= si.IsValid() && active && !activeMedium;
    if (activeSurface) {
      if (si.shape->HasBsdf()) resultTr = Spectrum(0);
      ray = si.SpawnRay(ray.D);
    }
    ray.MaxT = remainingDist;
    needsIntersection |= activeSurface;
    active &= (activeMedium || activeSurface) && !IsZero(resultTr);
    bool hasMedium = activeSurface && si.shape->HasMedia();
    if (hasMedium) medium = si.GetMedium(ray);
  }
  return {Mul(li, resultTr), dsr};
}

Spectrum VolPathLi(const Ray& ray0, const Scene& scene, Sampler& sampler, int maxDepth, int rrDepth) {
  Ray ray = ray0;
  const Medium* medium = scene.globalMedium;
  Spectrum throughput(1), result(0);
  Float eta = 1;
  int depth = 0;
  bool isSpecularPath = true;
  Interaction prevIts{};
  Float prevPdf = 1;
  uint32_t channel = std::min((uint32_t)(sampler.Next1D() * 3), 2u);  // vol_path.cpp:219-222
  SurfaceInteraction si{};
  for (;; depth++) {
    if (depth >= rrDepth) {
      Float rr = std::min(MaxCoeff(throughput) * Sqr(eta), Float(0.95));
      if (sampler.Next1D() > rr) break;
      throughput *= Rcp(rr);
    }
    if (maxDepth >= 0 && depth + 1 >= maxDepth) break;
    bool isSampleMedium = medium != nullptr;
    bool isSampleSurface = !isSampleMedium;
    bool isNullScatter = false, isMediumScatter = false, isMediumEscaped = false;
    bool isSpectral = isSampleMedium;
    bool notSpectral = false;
    MediumInteraction mei{};
    if (isSampleMedium) {
      mei = medium->SampleInteraction(ray, sampler.Next1D(), channel);
      if (medium->isHomogeneous && mei.IsValid()) ray.MaxT = mei.T;
      scene.RayIntersect(ray, si);
      if (si.T < mei.T) mei.T = Inf;
      if (isSpectral) {
        auto [tr, freeFlightPdf] = medium->EvalTrAndPdf(mei, si);
        Float pdf = freeFlightPdf[channel];
        Spectrum after = Mul(throughput, tr) / pdf;
        throughput = pdf <= 0 ? Spectrum(0) : after;
      }
      isMediumEscaped = isSampleMedium && !mei.IsValid();
      isSampleMedium &= mei.IsValid();
      bool nullScatter = sampler.Next1D() > mei.SigmaT[channel] / mei.CombinedExtinction[channel];
      isNullScatter |= nullScatter && isSampleMedium;
      isMediumScatter |= !isNullScatter && isSampleMedium;
      if (isSpectral && isNullScatter)
        throughput = Mul(throughput, mei.SigmaN) * mei.CombinedExtinction[channel] / mei.SigmaN[channel];
      if (isMediumScatter) prevIts = mei;
    }
    if (isNullScatter) {
      ray.O = mei.P;
      si.T = si.T - mei.T;
    }
    if (isMediumScatter) {
      if (isSpectral) throughput = Mul(throughput, mei.SigmaS) * mei.CombinedExtinction[channel] / mei.SigmaT[channel];
      if (notSpectral) throughput = Div(Mul(throughput, mei.SigmaS), mei.SigmaT);
      auto [li, dsr] = VolSampleLight(scene, sampler, medium, channel, mei);
      if (dsr.Pdf > 0) {
        Float phaseVal = medium->PhaseEval(mei, dsr.Dir);
        Float weight = dsr.IsDelta ? 1 : MisWeight(dsr.Pdf, phaseVal);
        result += Mul(throughput, li) * weight * phaseVal / dsr.Pdf;
      }
      auto [wo, phasePdf] = medium->PhaseSample(mei, sampler.Next2D());
      ray = mei.SpawnRay(wo);
      isSpecularPath = false;
      prevPdf = phasePdf;
    }
    isSampleSurface |= isMediumEscaped;
    if (isSampleSurface) {
      scene.RayIntersect(ray, si);
      const Light* hitLight = scene.GetLight(si);
      if (hitLight) {
        Float weight;
        if (isSpecularPath) {
          weight = 1;
        } else {
          Float lightPdf = scene.PdfLightDirection(hitLight, prevIts, si.ToDsr(prevIts));
          weight = MisWeight(prevPdf, lightPdf);
        }
        result += Mul(throughput, hitLight->Eval(si)) * weight;
      }
      if (!si.IsValid()) break;
      if (si.shape->IsLight()) break;
      if (!si.shape->HasBsdf()) {
        ray = si.SpawnRay(ray.D);
        medium = si.GetMedium(ray);
        depth--;
        continue;
      }
      // vol_path.cpp:28 drops the camera ray's differentials: textures see a zero footprint
      std::unique_ptr<Bsdf> resolved;
      const Bsdf* bsdf = scene.HitBsdf(si, resolved);
      if (bsdf->HasAnyTypeExceptDelta()) {
        auto [li, dsr] = VolSampleLight(scene, sampler, medium, channel, si);
        if (dsr.Pdf > 0) {
          V3 wo = si.ToLocal(dsr.Dir);
          Spectrum f = bsdf->Eval(si, wo);
          Float bsdfPdf = bsdf->Pdf(si, wo);
          Float weight = dsr.IsDelta ? 1 : MisWeight(dsr.Pdf, bsdfPdf);
          result += Mul(Mul(throughput, f), li) * weight / dsr.Pdf;
        }
      }
      Float lobe = sampler.Next1D();
      V2 dxi = sampler.Next2D();
      auto [bsr, f] = bsdf->Sample(si, lobe, dxi);
      if (bsr.Pdf <= 0) break;
      throughput = Mul(throughput, f) / bsr.Pdf;
      eta *= bsr.Eta;
      ray = si.SpawnRay(si.ToWorld(bsr.Wo));
      prevPdf = bsr.Pdf;
      prevIts = si;
      isSpecularPath = (bsr.TypeMask & BSDF_DELTA) != 0;
      medium = si.GetMedium(ray);
    }
    if (!(isSampleSurface | isSampleMedium)) break;
  }
  return result;
}

}  // namespace ref

// ======================================================================= C API
using namespace ref;

struct ref_scene {
  Scene scene;
};

static thread_local std::string g_err;

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

ref_scene* ref_scene_create(const rad_scene_desc* desc, int quirks) {
  try {
    auto* s = new ref_scene();
    s->scene.Create(*desc, quirks != 0);
    return s;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void ref_scene_destroy(ref_scene* s) { delete s; }

double ref_scene_build_ms(const ref_scene* s) { return s->scene.accel.buildMs; }

void ref_scene_bounds(const ref_scene* s, float* lo, float* hi) {
  const Bounds& b = s->scene.accel.world;
  for (int a = 0; a < 3; a++) {
    lo[a] = b.lo[a];
    hi[a] = b.hi[a];
  }
}

int ref_intersect_closest(const ref_scene* s, const rad_ray* rays, uint32_t n, int brute, rad_hit* hits) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    Ray r{V3(rays[i].o[0], rays[i].o[1], rays[i].o[2]), V3(rays[i].d[0], rays[i].d[1], rays[i].d[2]), rays[i].tmin,
          rays[i].tmax};
    HitShapeRecord rec;
    bool h = brute == 2 ? s->scene.accel.ClosestCulled(r, rec)
             : brute    ? s->scene.accel.ClosestBrute(r, rec)
                        : s->scene.accel.Closest(r, rec);
    hits[i].t = h ? rec.T : r.MaxT;
    hits[i].u = h ? rec.PrimitiveUV.x : 0;
    hits[i].v = h ? rec.PrimitiveUV.y : 0;
    hits[i].prim = h ? rec.PrimitiveIndex : 0xffffffffu;
    hits[i].geom = h ? rec.ShapeIndex : 0xffffffffu;
  }
  return 0;
}

int ref_intersect_any(const ref_scene* s, const rad_ray* rays, uint32_t n, int brute, uint8_t* occ) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    Ray r{V3(rays[i].o[0], rays[i].o[1], rays[i].o[2]), V3(rays[i].d[0], rays[i].d[1], rays[i].d[2]), rays[i].tmin,
          rays[i].tmax};
    occ[i] = (brute == 2 ? s->scene.accel.OccludedCulled(r)
              : brute    ? s->scene.accel.OccludedBrute(r)
                         : s->scene.accel.Occluded(r))
                 ? 1
                 : 0;
  }
  return 0;
}

// t of one ray against one primitive with the full [tmin, tmax] window; -1 if no hit.
float ref_intersect_prim(const ref_scene* s, const rad_ray* ray, uint32_t geom, uint32_t prim) {
  Ray r{V3(ray->o[0], ray->o[1], ray->o[2]), V3(ray->d[0], ray->d[1], ray->d[2]), ray->tmin, ray->tmax};
  HitShapeRecord rec;
  if (geom >= s->scene.shapes.size()) return -1;
  if (!s->scene.shapes[geom]->IntersectPrim(prim, r, r.MinT, r.MaxT, rec)) return -1;
  return rec.T;
}

void ref_camera_ray(const ref_scene* s, float sx, float sy, rad_ray* out) {
  Ray r = s->scene.camera.SampleRay(V2(sx, sy));
  out->o[0] = r.O.x; out->o[1] = r.O.y; out->o[2] = r.O.z;
  out->d[0] = r.D.x; out->d[1] = r.D.y; out->d[2] = r.D.z;
  out->tmin = r.MinT;
  out->tmax = r.MaxT;
}

// SampleRenderer::Start (sample_renderer.cpp:60-122) over fixed tiles with
// OpenMP threads in place of TBB.  Writes the mean (sum * 1/spp) and, when
// requested, the per-pixel sum of squares of accepted samples.
int ref_render(const ref_scene* s, const ref_render_params* p, float* fb, float* sumsq, double* seconds) {
  const Scene& scene = s->scene;
  const int W = scene.camera.width, H = scene.camera.height;
  const int tile = p->tile_size > 0 ? p->tile_size : 16;
  const int x0 = p->x_begin, x1 = p->x_end > 0 ? p->x_end : W;
  const int y0 = p->y_begin, y1 = p->y_end > 0 ? p->y_end : H;
  const int tilesX = (x1 - x0 + tile - 1) / tile, tilesY = (y1 - y0 + tile - 1) / tile;
  const uint32_t spp = p->spp;
  const bool quirks = scene.quirks;
  int threads = p->threads > 0 ? p->threads : omp_get_max_threads();
  auto t0 = std::chrono::steady_clock::now();
  std::atomic<int> err{0};
  const int step = p->tile_step > 1 ? p->tile_step : 1;
  const int ntiles = (tilesX * tilesY + step - 1) / step;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int ti = 0; ti < ntiles; ti++) {
    const int t = ti * step;
    int bx = x0 + (t % tilesX) * tile, by = y0 + (t / tilesX) * tile;
    int ex = std::min(bx + tile, x1), ey = std::min(by + tile, y1);
    uint32_t seed = uint32_t(bx) * uint32_t(W) + uint32_t(by);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<Float> dist;
    Sampler sampler(p->seed + seed);
    for (int y = by; y < ey; y++) {
      for (int x = bx; x < ex; x++) {
        Spectrum acc(0), acc2(0);
        for (uint32_t i = 0; i < spp; i++) {
          Float jx = dist(rng);
          Float jy = dist(rng);
          RayDiff rd;
          Ray ray = scene.camera.SampleRayDifferential(V2(x + jx, y + jy), rd);
          Spectrum li = p->integrator == RAD_INTEGRATOR_VOL_PATH ? VolPathLi(ray, scene, sampler, p->max_depth, p->rr_depth)
                                                                 : PathLi(ray, scene, sampler, p->max_depth, p->rr_depth, rd);
          bool bad = HasNaN(li);
          if (quirks) {  // spectrum.h:30-31 checks the first channel only
            bad = bad || std::isinf(li.x) || li.x < 0;
          } else {
            bad = bad || std::isinf(li.x) || std::isinf(li.y) || std::isinf(li.z) || li.x < 0 || li.y < 0 || li.z < 0;
          }
          if (!bad) {
            acc += li;
            acc2 += Mul(li, li);
          }
        }
        size_t o = (size_t(y) * W + x) * 3;
        Float coeff = 1.0f / spp;
        fb[o + 0] = acc.x * coeff;
        fb[o + 1] = acc.y * coeff;
        fb[o + 2] = acc.z * coeff;
        if (sumsq) {
          sumsq[o + 0] = acc2.x;
          sumsq[o + 1] = acc2.y;
          sumsq[o + 2] = acc2.z;
        }
      }
    }
  }
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return err.load();
}

// Per-function probes: BSDF eval / pdf / sample in shading space.
int ref_bsdf_eval(const rad_bsdf_desc* d, uint32_t count, int quirks, const float* in, uint32_t n, float* out) {
  try {
    auto b = MakeBsdf(d, count, 0, quirks != 0);
    for (uint32_t i = 0; i < n; i++) {
      SurfaceInteraction si{};
      si.Wi = V3(in[6 * i], in[6 * i + 1], in[6 * i + 2]);
      V3 wo(in[6 * i + 3], in[6 * i + 4], in[6 * i + 5]);
      Spectrum f = b->Eval(si, wo);
      out[4 * i] = f.x;
      out[4 * i + 1] = f.y;
      out[4 * i + 2] = f.z;
      out[4 * i + 3] = b->Pdf(si, wo);
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int ref_bsdf_sample(const rad_bsdf_desc* d, uint32_t count, int quirks, const float* in, uint32_t n, float* out) {
  try {
    auto b = MakeBsdf(d, count, 0, quirks != 0);
    for (uint32_t i = 0; i < n; i++) {
      SurfaceInteraction si{};
      si.Wi = V3(in[6 * i], in[6 * i + 1], in[6 * i + 2]);
      auto [bsr, f] = b->Sample(si, in[6 * i + 3], V2(in[6 * i + 4], in[6 * i + 5]));
      float* o = out + 8 * i;
      o[0] = bsr.Wo.x; o[1] = bsr.Wo.y; o[2] = bsr.Wo.z;
      o[3] = bsr.Pdf;
      o[4] = f.x; o[5] = f.y; o[6] = f.z;
      o[7] = bsr.Eta;
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

float ref_erfinv(float a) { return ref::RefErfInv(a); }

int ref_texture_eval(const rad_texture_desc* t, uint32_t count, uint32_t index, const float* in, uint32_t n, float* out) {
  try {
    std::vector<std::unique_ptr<Texture>> tex;
    BuildTextures(t, count, tex);
    if (index >= count) throw std::runtime_error("texture index out of range");
    for (uint32_t i = 0; i < n; i++) {
      const float* q = in + 6 * i;
      V3 v = tex[index]->Eval(V2(q[0], q[1]), V2(q[2], q[3]), V2(q[4], q[5]));
      out[3 * i] = v.x;
      out[3 * i + 1] = v.y;
      out[3 * i + 2] = v.z;
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Texture coordinates and their screen-space derivatives at the first hit of
// the camera ray through each screen position (0 where the ray misses).
int ref_camera_footprint(const ref_scene* s, const float* in, uint32_t n, float* out) {
  const Scene& scene = s->scene;
  for (uint32_t i = 0; i < n; i++) {
    RayDiff rd;
    Ray ray = scene.camera.SampleRayDifferential(V2(in[2 * i], in[2 * i + 1]), rd);
    SurfaceInteraction si{};
    float* o = out + 6 * i;
    for (int k = 0; k < 6; k++) o[k] = 0.f;
    if (!scene.RayIntersect(ray, si) || !si.shape) continue;
    ComputeUVDifferentials(si, rd);
    o[0] = si.UV.x, o[1] = si.UV.y, o[2] = si.dUVdX.x, o[3] = si.dUVdX.y, o[4] = si.dUVdY.x, o[5] = si.dUVdY.y;
  }
  return 0;
}

int ref_grid_eval(const rad_medium_desc* d, const float* in, uint32_t n, float* out) {
  GridVolume g;
  g.sx = d->grid_size[0];
  g.sy = d->grid_size[1];
  g.sz = d->grid_size[2];
  g.channels = d->grid_channels;
  g.wrap = d->grid_wrap;
  g.data.assign(d->grid, d->grid + size_t(g.sx) * g.sy * g.sz * g.channels);
  for (uint32_t i = 0; i < n; i++) {
    Spectrum v = g.Eval(V3(in[3 * i], in[3 * i + 1], in[3 * i + 2]));
    out[3 * i] = v.x;
    out[3 * i + 1] = v.y;
    out[3 * i + 2] = v.z;
  }
  return 0;
}

// Delta-tracking transmittance estimate along one ray through one medium
// (tests compare its mean with quadrature of the grid).
float ref_medium_transmittance(const ref_scene* s, int medium, const rad_ray* ray, uint32_t channel, uint32_t seed) {
  const Medium* m = s->scene.media[medium].get();
  Sampler sampler(seed);
  Ray r{V3(ray->o[0], ray->o[1], ray->o[2]), V3(ray->d[0], ray->d[1], ray->d[2]), ray->tmin, ray->tmax};
  // Ratio-free delta tracking: survive while collisions are null.
  for (int guard = 0; guard < 100000; guard++) {
    MediumInteraction mei = m->SampleInteraction(r, sampler.Next1D(), channel);
    if (!mei.IsValid()) return 1.f;
    if (sampler.Next1D() < mei.SigmaT[channel] / mei.CombinedExtinction[channel]) return 0.f;
    r.MaxT = r.MaxT - mei.T;
    r.O = mei.P;
    r.MinT = 0;
  }
  return 0.f;
}

int ref_phase_eval(int type, float g, const float* in, uint32_t n, float* out) {
  HomogeneousMedium m;
  m.phaseType = type;
  m.g = g;
  for (uint32_t i = 0; i < n; i++) {
    MediumInteraction mi{};
    mi.Wi = V3(in[6 * i], in[6 * i + 1], in[6 * i + 2]);
    mi.Shading = Frame(mi.Wi);
    out[i] = m.PhaseEval(mi, V3(in[6 * i + 3], in[6 * i + 4], in[6 * i + 5]));
  }
  return 0;
}

int ref_phase_sample(int type, float g, const float* in, uint32_t n, float* out) {
  HomogeneousMedium m;
  m.phaseType = type;
  m.g = g;
  for (uint32_t i = 0; i < n; i++) {
    MediumInteraction mi{};
    mi.Wi = V3(in[5 * i], in[5 * i + 1], in[5 * i + 2]);
    mi.Shading = Frame(mi.Wi);
    auto [wo, pdf] = m.PhaseSample(mi, V2(in[5 * i + 3], in[5 * i + 4]));
    float* o = out + 4 * i;
    o[0] = wo.x; o[1] = wo.y; o[2] = wo.z;
    o[3] = pdf;
  }
  return 0;
}

int ref_light_sample(const ref_scene* s, uint32_t light, const float* in, uint32_t n, float* out) {
  if (light >= s->scene.lights.size()) {
    g_err = "light index out of range";
    return 1;
  }
  const Light* L = s->scene.lights[light].get();
  for (uint32_t i = 0; i < n; i++) {
    Interaction ref{};
    ref.P = V3(in[5 * i], in[5 * i + 1], in[5 * i + 2]);
    auto [dsr, li] = L->SampleDirection(ref, V2(in[5 * i + 3], in[5 * i + 4]));
    float* o = out + 15 * i;
    o[0] = dsr.Dir.x; o[1] = dsr.Dir.y; o[2] = dsr.Dir.z;
    o[3] = dsr.Dist;
    o[4] = dsr.Pdf;
    o[5] = dsr.P.x; o[6] = dsr.P.y; o[7] = dsr.P.z;
    o[8] = dsr.N.x; o[9] = dsr.N.y; o[10] = dsr.N.z;
    o[11] = li.x; o[12] = li.y; o[13] = li.z;
    o[14] = L->PdfDirection(ref, dsr);
  }
  return 0;
}

}  // extern "C"
