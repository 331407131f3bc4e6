// ORACLE / TEST INFRASTRUCTURE ONLY (see ref_math.h).
// BSDFs of the restatement: diffuse, rough_metal, rough_glass, disney,
// plastic, rough_plastic, perfect_mirror, perfect_glass and the two_side /
// mask wrappers, with the GGX / Beckmann microfacet distributions and the
// Fresnel terms.
#include "ref_scene.h"

namespace ref {

// ---------------------------------------------------------------- Fresnel (fresnel.cpp:10-77)
static V3 Reflect(const V3& wi, const V3& wh) { return Fmadd(wh, V3(2 * Dot(wi, wh)), -wi); }
static V3 Refract(const V3& wi, const V3& wh, Float cosThetaT, Float etaTI) {
  return Fmadd(wh, V3(Fmadd(Dot(wi, wh), etaTI, cosThetaT)), -(wi * etaTI));
}
struct DielectricF {
  Float r, cosThetaT, etaIT, etaTI;
};
static DielectricF FresnelDielectric(Float cosThetaI, Float eta) {
  bool isOutside = cosThetaI >= 0;
  Float rcpEta = Rcp(eta);
  Float etaIT = isOutside ? eta : rcpEta;
  Float etaTI = isOutside ? rcpEta : eta;
  Float sqrCosThetaT = Fmadd(-Fmadd(-cosThetaI, cosThetaI, 1), etaTI * etaTI, 1);
  Float absCosThetaI = std::abs(cosThetaI);
  Float absCosThetaT = SafeSqrt(sqrCosThetaT);
  bool indexMatched = eta == 1;
  bool specialCase = indexMatched || absCosThetaI == 0;
  Float rSC = indexMatched ? 0.f : 1.f;
  Float aS = Fmadd(-etaIT, absCosThetaT, absCosThetaI) / Fmadd(etaIT, absCosThetaT, absCosThetaI);
  Float aP = Fmadd(-etaIT, absCosThetaI, absCosThetaT) / Fmadd(etaIT, absCosThetaI, absCosThetaT);
  Float r = Float(0.5) * (Sqr(aS) + Sqr(aP));
  if (specialCase) r = rSC;
  Float cosThetaT = cosThetaI >= 0 ? -absCosThetaT : absCosThetaT;
  return {r, cosThetaT, etaIT, etaTI};
}
static Spectrum FresnelConductor(Float cosThetaI, const Spectrum& eta, const Spectrum& k) {
  Float cosThetaI2 = cosThetaI * cosThetaI, sinThetaI2 = 1 - cosThetaI2, sinThetaI4 = sinThetaI2 * sinThetaI2;
  Spectrum out;
  for (int c = 0; c < 3; c++) {
    Float etaR = eta[c], etaI = k[c];
    Float temp1 = etaR * etaR - etaI * etaI - sinThetaI2;
    Float a2pb2 = std::sqrt(std::max(temp1 * temp1 + etaI * etaI * etaR * etaR * 4, Float(0)));
    Float a = std::sqrt(std::max((a2pb2 + temp1) * Float(0.5), Float(0)));
    Float term1 = a2pb2 + cosThetaI2;
    Float term2 = a * 2 * cosThetaI;
    Float rs = (term1 - term2) * (1 / (term1 + term2));
    Float term3 = a2pb2 * cosThetaI2 + sinThetaI4;
    Float term4 = term2 * sinThetaI2;
    Float rp = rs * (term3 - term4) * (1 / (term3 + term4));
    out[c] = (rs + rp) * Float(0.5);
  }
  return out;
}

// ---------------------------------------------------------------- microfacet (microfacet.cpp)
// erf^-1 for the Beckmann visible-normal sampler, from the mathematics rather
// than the reference's polynomial: the argument enters through 1 - a^2
// clamped at FLT_MIN as in math_ext.h:114-140 (so |a| >= 1 saturates), and
// y = erfc^-1(c) with c = (1 - a^2) / (1 + sqrt(a^2)) is found by Newton on
// log(erfc(y)) = log(c) in double precision.
static Float ErfInv(Float a) {
  if (std::abs(a) < Float(0.5)) {  // well-conditioned core: Newton on erf(y) = |a|
    double A = std::abs(a), y = A * 0.886226925452758;
    for (int i = 0; i < 30; i++) {
      double step = (std::erf(y) - A) / (1.1283791670955126 * std::exp(-y * y));
      y -= step;
      if (std::abs(step) <= 1e-16 * std::max(1e-300, y)) break;
    }
    return std::copysign(Float(y), a);
  }
  Float q = std::max(std::fma(a, -a, Float(1)), std::numeric_limits<Float>::min());
  double Q = q;
  double c = Q / (1.0 + std::sqrt(1.0 - Q));
  if (c >= 1.0) return Float(0) * a;
  double L = std::log(c);
  double y = c > 0.5 ? (1.0 - c) * 0.886226925452758 : std::sqrt(std::max(-L - 0.5 * std::log(-L), 1e-6));
  for (int i = 0; i < 60; i++) {
    double e = std::erfc(y);
    double f = std::log(e) - L;
    double fp = -1.1283791670955126 * std::exp(-y * y) / e;
    double step = f / fp;
    y -= step;
    if (y < 0.0) y = 0.0;
    if (std::abs(step) <= 1e-15 * std::max(1.0, y)) break;
  }
  return std::copysign(Float(y), a);
}

struct Microfacet {
  bool beckmann = false;
  Float alphaX = 0.1f, alphaY = 0.1f;
  bool sampleVisible = true;
  bool quirks = true;

  Float D(const V3& wh) const {
    Float alphaXY = alphaX * alphaY;
    Float cosTheta = Frame::CosTheta(wh);
    if (beckmann) {
      Float cosTheta2 = Sqr(cosTheta);
      Float result = std::exp(-(Sqr(wh.x / alphaX) + Sqr(wh.y / alphaY)) / cosTheta2) / (PI * alphaXY * Sqr(cosTheta2));
      return result * cosTheta > 1e-20f ? result : 0;
    }
    Float result = Rcp(PI * alphaXY * Sqr(Sqr(wh.x / alphaX) + Sqr(wh.y / alphaY) + Sqr(wh.z)));
    return result * cosTheta > 1e-20f ? result : 0;
  }
  Float SmithG1(const V3& v, const V3& wh) const {
    Float xyAlpha2 = Sqr(alphaX * v.x) + Sqr(alphaY * v.y);
    Float tanThetaAlpha2 = xyAlpha2 / Sqr(v.z);
    Float result;
    if (beckmann) {
      Float a = Rsqrt(tanThetaAlpha2);
      Float aSqr = Sqr(a);
      result = a >= 1.6f ? 1.f : (3.535f * a + 2.181f * aSqr) / (1.f + 2.276f * a + 2.577f * aSqr);
    } else {
      result = 2 / (1 + std::sqrt(1 + tanThetaAlpha2));
    }
    if (xyAlpha2 == 0) result = 1;
    if (Dot(v, wh) * Frame::CosTheta(v) <= 0) result = 0;
    return result;
  }
  Float G(const V3& wi, const V3& wo, const V3& wh) const { return SmithG1(wi, wh) * SmithG1(wo, wh); }
  Float Pdf(const V3& wi, const V3& wh) const {
    Float result = D(wh);
    if (sampleVisible) {
      result *= SmithG1(wi, wh) * AbsDot(wi, wh) / Frame::CosTheta(wi);
    } else {
      result *= Frame::CosTheta(wh);
    }
    return result;
  }
  static V2 SampleVisibleGGX(Float cosThetaI, V2 sample) {  // microfacet.cpp:202-236
    if (cosThetaI > 0.9999) {
      Float r = std::sqrt(sample.x / (1 - sample.x));
      Float phi = Float(6.28318530718) * sample.y;
      return V2(r * std::cos(phi), r * std::sin(phi));
    }
    Float sinTheta = std::sqrt(std::max((Float)0, (Float)1 - Sqr(cosThetaI)));
    Float tanTheta = sinTheta / cosThetaI;
    Float a = 1 / tanTheta;
    Float G1 = 2 / (1 + std::sqrt(1 + 1 / Sqr(a)));
    Float A = 2 * sample.x / G1 - 1;
    Float tmp = 1.f / Fmadd(A, A, -1);
    if (tmp > 1e10) tmp = Float(1e10);
    Float B = tanTheta;
    Float Dd = std::sqrt(std::max(Float(Sqr(B) * Sqr(tmp) - (Sqr(A) - Sqr(B)) * tmp), Float(0)));
    Float slopeX1 = Fmadd(B, tmp, -Dd);
    Float slopeX2 = Fmadd(B, tmp, Dd);
    Float slopeX = (A < 0 || slopeX2 > 1 / tanTheta) ? slopeX1 : slopeX2;
    Float S;
    if (sample.y > 0.5f) {
      S = 1;
      sample.y = 2 * (sample.y - Float(0.5));
    } else {
      S = -1;
      sample.y = 2 * (Float(0.5) - sample.y);
    }
    Float z = (sample.y * (sample.y * (sample.y * 0.27385f - 0.73369f) + 0.46341f)) /
              (sample.y * (sample.y * (sample.y * 0.093073f + 0.309420f) - 1) + 0.597999f);
    Float slopeY = S * z * std::sqrt(1 + slopeX * slopeX);
    return V2(slopeX, slopeY);
  }
  static V2 SampleVisibleBeckmann(Float cosThetaI, V2 sample) {  // microfacet.cpp:82-97
    Float tanThetaI = SafeSqrt(Fmadd(-cosThetaI, cosThetaI, Float(1))) / cosThetaI;
    Float cotThetaI = Rcp(tanThetaI);
    Float maxval = std::erf(cotThetaI);
    sample.x = std::max(std::min(sample.x, Float(1 - 1e-6)), Float(1e-6));
    sample.y = std::max(std::min(sample.y, Float(1 - 1e-6)), Float(1e-6));
    Float x = maxval - (maxval + 1) * std::erf(std::sqrt(-std::log(sample.x)));
    sample.x *= 1 + maxval + (1 / std::sqrt(PI)) * tanThetaI * std::exp(-Sqr(cotThetaI));
    for (int i = 0; i < 3; i++) {
      Float slope = ErfInv(x);
      Float value = 1 + x + (1 / std::sqrt(PI)) * tanThetaI * std::exp(-Sqr(slope)) - sample.x;
      Float derivative = 1 - slope * tanThetaI;
      x -= value / derivative;
    }
    return V2(ErfInv(x), ErfInv(Fmadd(Float(2), sample.y, Float(-1))));
  }
  std::pair<V3, Float> Sample(const V3& wi, const V2& xi) const {
    if (sampleVisible) {
      V3 wiP = Normalized(V3(alphaX * wi.x, alphaY * wi.y, wi.z));
      auto [sinPhi, cosPhi] = Frame::SinCosPhi(wiP, quirks);
      Float cosTheta = Frame::CosTheta(wiP);
      V2 slope = beckmann ? SampleVisibleBeckmann(cosTheta, xi) : SampleVisibleGGX(cosTheta, xi);
      slope = V2(Fmadd(cosPhi, slope.x, -(sinPhi * slope.y)) * alphaX, Fmadd(sinPhi, slope.x, cosPhi * slope.y) * alphaY);
      V3 m = Normalized(V3(-slope.x, -slope.y, 1));
      Float pdf = D(m) * SmithG1(wi, m) * AbsDot(wi, m) / Frame::CosTheta(wi);
      return {m, pdf};
    }
    Float sinPhi, cosPhi, alpha2;
    if (alphaX == alphaY) {
      sinPhi = std::sin((2 * PI) * xi.y);
      cosPhi = std::cos((2 * PI) * xi.y);
      alpha2 = alphaX * alphaY;
    } else {
      Float ratio = alphaY / alphaX;
      Float tmp = ratio * std::tan((2 * PI) * xi.y);
      cosPhi = Rsqrt(Fmadd(tmp, tmp, 1));
      cosPhi = MulSign(cosPhi, std::abs(xi.y - Float(0.5)) - Float(0.25));
      sinPhi = cosPhi * tmp;
      alpha2 = Rcp(Sqr(cosPhi / alphaX) + Sqr(sinPhi / alphaY));
    }
    if (beckmann) {
      Float cosTheta = Rsqrt(Fmadd(-alpha2, std::log(1 - xi.x), 1));
      Float cos2Theta = Sqr(cosTheta);
      Float cos3Theta = std::max(cos2Theta * cosTheta, Float(1e-20));
      Float pdf = (1.f - xi.x) / (PI * alphaX * alphaY * cos3Theta);
      Float sinTheta = std::sqrt(1 - cos2Theta);
      return {V3(cosPhi * sinTheta, sinPhi * sinTheta, cosTheta), pdf};
    }
    Float tanThetaM2 = alpha2 * xi.x / (1.f - xi.x);
    Float cosTheta = Rsqrt(1.f + tanThetaM2);
    Float cos2Theta = Sqr(cosTheta);
    Float temp = 1 + tanThetaM2 / alpha2;
    Float cos3Theta = std::max(cos2Theta * cosTheta, Float(1e-20));
    Float pdf = Rcp(PI * alphaX * alphaY * cos3Theta * Sqr(temp));
    Float sinTheta = std::sqrt(1 - cos2Theta);
    return {V3(cosPhi * sinTheta, sinPhi * sinTheta, cosTheta), pdf};
  }
};

// ---------------------------------------------------------------- diffuse (bsdf/diffuse.cpp:40-99)
struct Diffuse final : Bsdf {
  Spectrum reflectance;
  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float, const V2& dirXi) const override {
    if (Frame::CosTheta(si.Wi) <= 0) return {{}, Spectrum(0)};
    BsdfSampleResult bsr{};
    bsr.Wo = Warp::SquareToCosineHemisphere(dirXi);
    bsr.Pdf = Warp::SquareToCosineHemispherePdf(bsr.Wo);
    bsr.Eta = 1;
    bsr.TypeMask = flags;
    return {bsr, reflectance * (1 / PI) * Frame::CosTheta(bsr.Wo)};
  }
  Spectrum Eval(const SurfaceInteraction& si, const V3& wo) const override {
    if (Frame::CosTheta(si.Wi) <= 0 || Frame::CosTheta(wo) <= 0) return Spectrum(0);
    return reflectance * (1 / PI) * Frame::CosTheta(wo);
  }
  Float Pdf(const SurfaceInteraction& si, const V3& wo) const override {
    if (Frame::CosTheta(si.Wi) <= 0 || Frame::CosTheta(wo) <= 0) return 0;
    return Warp::SquareToCosineHemispherePdf(wo);
  }
};

// ---------------------------------------------------------------- rough_metal (bsdf/rough_metal.cpp:20-176)
struct RoughMetal final : Bsdf {
  Spectrum reflectance, eta, k;
  Microfacet dist;
  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float, const V2& dirXi) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi);
    if (cosThetaI <= 0) return {{}, Spectrum(0)};
    BsdfSampleResult bsr{};
    auto [m, pdf] = dist.Sample(si.Wi, dirXi);
    bsr.Pdf = pdf;
    bsr.Wo = Reflect(si.Wi, m);
    bsr.Eta = 1;
    bsr.TypeMask = flags;
    if (bsr.Pdf <= 0 || Frame::CosTheta(bsr.Wo) <= 0) return {{}, Spectrum(0)};
    bsr.Pdf /= 4 * Dot(bsr.Wo, m);
    Float D = dist.D(m);
    Float G = dist.G(si.Wi, bsr.Wo, m);
    Spectrum F = FresnelConductor(Dot(si.Wi, m), eta, k);
    Spectrum brdf = (F * D * G) / (4 * Frame::CosTheta(si.Wi) * Frame::CosTheta(bsr.Wo));
    return {bsr, Mul(reflectance, brdf) * Frame::CosTheta(bsr.Wo)};
  }
  Spectrum Eval(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi), cosThetaO = Frame::CosTheta(wo);
    if (cosThetaI <= 0 || cosThetaO <= 0) return Spectrum(0);
    V3 wh = Normalized(wo + si.Wi);
    Spectrum F = FresnelConductor(Dot(si.Wi, wh), eta, k);
    Float D = dist.D(wh);
    Float G = dist.G(si.Wi, wo, wh);
    Spectrum brdf = Abs(F * D * G) / (cosThetaI * cosThetaO * 4);
    return Mul(reflectance, brdf) * cosThetaO;
  }
  Float Pdf(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi), cosThetaO = Frame::CosTheta(wo);
    if (cosThetaI <= 0 || cosThetaO <= 0) return 0;
    V3 wh = Normalized(wo + si.Wi);
    return dist.Pdf(si.Wi, wh) / (4 * Dot(wo, wh));
  }
};

// ---------------------------------------------------------------- rough_glass (bsdf/rough_glass.cpp:22-221)
struct RoughGlass final : Bsdf {
  Spectrum reflectance, transmittance;
  Float eta = 1.5f, invEta = 1 / 1.5f;
  Microfacet dist;
  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float lobeXi, const V2& dirXi) const override {
    BsdfSampleResult bsr{};
    Float cosThetaI = Frame::CosTheta(si.Wi);
    if (cosThetaI == 0) return {{}, Spectrum(0)};
    auto [wh, pdf] = dist.Sample(MulSign(si.Wi, cosThetaI), dirXi);
    if (pdf == 0) return {{}, Spectrum(0)};
    DielectricF f = FresnelDielectric(Dot(si.Wi, wh), eta);
    bool selectedR = lobeXi <= f.r;
    bsr.Wo = selectedR ? Reflect(si.Wi, wh) : Refract(si.Wi, wh, f.cosThetaT, f.etaTI);
    bool isReflect = Frame::IsSameHemisphere(si.Wi, bsr.Wo);
    bsr.Eta = isReflect ? 1 : f.etaIT;
    bsr.TypeMask = BSDF_GLOSSY | (isReflect ? BSDF_REFLECTION : BSDF_TRANSMISSION);
    bsr.Pdf = Pdf(si, bsr.Wo);
    if (bsr.Pdf <= 0) return {{}, Spectrum(0)};
    return {bsr, Eval(si, bsr.Wo)};
  }
  Spectrum Eval(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi), cosThetaO = Frame::CosTheta(wo);
    if (cosThetaI == 0) return Spectrum(0);
    bool reflect = cosThetaI * cosThetaO > 0;
    Float e = cosThetaI > 0 ? eta : invEta;
    Float ie = cosThetaI > 0 ? invEta : eta;
    V3 m = Normalized(si.Wi + wo * (reflect ? 1 : e));
    m = MulSign(m, Frame::CosTheta(m));
    Float D = dist.D(m);
    Float F = FresnelDielectric(Dot(si.Wi, m), eta).r;
    Float G = dist.G(si.Wi, wo, m);
    Spectrum result(0);
    if (reflect) {
      Float brdf = std::abs(F * D * G / (cosThetaI * 4));
      result = reflectance * brdf;
    }
    if (!reflect) {
      Float scale = Sqr(ie);  // TransportMode::Radiance
      Float idm = Dot(si.Wi, m), odm = Dot(wo, m);
      Float btdf = std::abs((scale * (1 - F) * D * G * e * e * idm * odm) / (cosThetaI * Sqr(idm + e * odm)));
      result = transmittance * btdf;
    }
    return result;
  }
  Float Pdf(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi), cosThetaO = Frame::CosTheta(wo);
    if (cosThetaI == 0) return 0;
    bool reflect = cosThetaI * cosThetaO > 0;
    Float e = cosThetaI > 0 ? eta : invEta;
    V3 m = Normalized(si.Wi + wo * (reflect ? 1 : e));
    m = MulSign(m, Frame::CosTheta(m));
    Float idm = Dot(si.Wi, m), odm = Dot(wo, m);
    if (idm * cosThetaI <= 0 || odm * cosThetaO <= 0) return 0;
    Float dwhdwo = reflect ? Rcp(odm * 4) : ((e * e * odm) / (Sqr(idm + e * odm)));
    Float pdf = dist.Pdf(MulSign(si.Wi, cosThetaI), m);
    Float F = FresnelDielectric(idm, eta).r;
    pdf *= reflect ? F : (1 - F);
    return std::abs(pdf * dwhdwo);
  }
};

// ---------------------------------------------------------------- disney (bsdf/disney.cpp:28-613)
struct Disney final : Bsdf {
  Spectrum baseColor;
  Float roughness = 0.5f, anisotropic = 0, specTrans = 0, sheen = 0, sheenTint = 0, flatness = 0, specTint = 0,
        metallic = 0, clearcoat = 0, clearcoatGloss = 0;
  Float eta = 1.5f;
  Float specularSample = 1, clearcoatSample = 1, diffuseReflectSample = 1;
  bool hasClearcoat = false, hasSheen = false, hasSpecTrans = false, hasMetallic = false, hasSpecTint = false,
       hasSheenTint = false, hasAnisotropic = false, hasFlatness = false;

  struct GTR1 {
    Float alpha;
    V3 Sample(const V2& xi) const {
      Float a = (Float(2) * PI) * xi.x;
      Float sinPhi = std::sin(a), cosPhi = std::cos(a);
      Float alpha2 = Sqr(alpha);
      Float cosTheta2 = (Float(1) - std::pow(alpha2, Float(1) - xi.y)) / (Float(1) - alpha2);
      Float sinTheta = std::sqrt(std::max(Float(0), Float(1) - cosTheta2));
      Float cosTheta = std::sqrt(std::max(Float(0), cosTheta2));
      return V3(cosPhi * sinTheta, sinPhi * sinTheta, cosTheta);
    }
    Float Eval(const V3& m) const {
      Float cosTheta = Frame::CosTheta(m);
      Float cosTheta2 = Sqr(cosTheta), alpha2 = Sqr(alpha);
      Float result = (alpha2 - Float(1)) / (PI * std::log(alpha2) * (Float(1) + (alpha2 - Float(1)) * cosTheta2));
      return result * cosTheta > Float(1e-20f) ? result : Float(0);
    }
    Float Pdf(const V3& m) const { return m.z < 0 ? 0 : Frame::CosTheta(m) * Eval(m); }
  };

  std::pair<Float, Float> RoughnessToAlpha() const {
    Float r2 = Sqr(roughness);
    if (!hasAnisotropic) {
      Float a = std::max(Float(0.001), r2);
      return {a, a};
    }
    Float aspect = std::sqrt(Float(1) - Float(0.9) * anisotropic);
    return {std::max(Float(0.001), r2 / aspect), std::max(Float(0.001), r2 * aspect)};
  }
  Microfacet SpecDist() const {
    auto [ax, ay] = RoughnessToAlpha();
    Microfacet d;
    d.alphaX = ax;
    d.alphaY = ay;
    d.sampleVisible = true;
    d.quirks = quirks;
    return d;
  }
  static bool CheckMacMicNormal(const V3& m, const V3& wi, const V3& wo, Float cosThetaI, bool reflection) {
    if (reflection) return Dot(wi, MulSign(m, cosThetaI)) > 0 && Dot(wo, MulSign(m, cosThetaI)) > 0;
    return Dot(wi, MulSign(m, cosThetaI)) > 0 && Dot(wo, cosThetaI >= 0 ? -m : m) > 0;
  }
  static Float SchlickWeight(Float cosI) {
    Float m = std::min(std::max(Float(1) - cosI, Float(0)), Float(1));
    return Sqr(Sqr(m)) * m;
  }
  static Float SchlickF(Float R0, Float cosThetaI, Float eta) {
    bool isOutside = cosThetaI >= 0;
    Float rcpEta = Rcp(eta);
    Float etaIT = isOutside ? eta : rcpEta, etaTI = isOutside ? rcpEta : eta;
    Float sqrCosThetaT = Fmadd(-Fmadd(-cosThetaI, cosThetaI, Float(1)), Sqr(etaTI), Float(1));
    Float cosThetaT = SafeSqrt(sqrCosThetaT);
    return etaIT > 1 ? Lerp(SchlickWeight(std::abs(cosThetaI)), Float(1), R0) : Lerp(SchlickWeight(cosThetaT), Float(1), R0);
  }
  static Spectrum SchlickFS(const Spectrum& R0, Float cosThetaI, Float eta) {
    bool isOutside = cosThetaI >= 0;
    Float rcpEta = Rcp(eta);
    Float etaIT = isOutside ? eta : rcpEta, etaTI = isOutside ? rcpEta : eta;
    Float sqrCosThetaT = Fmadd(-Fmadd(-cosThetaI, cosThetaI, Float(1)), Sqr(etaTI), Float(1));
    Float cosThetaT = SafeSqrt(sqrCosThetaT);
    Float w = etaIT > 1 ? SchlickWeight(std::abs(cosThetaI)) : SchlickWeight(cosThetaT);
    // LerpSpectrum(a, b, t) = a + t * (b - a) with a = w, b = 1, t = R0 (spectrum.h:59-61)
    return Spectrum(w + R0.x * (1 - w), w + R0.y * (1 - w), w + R0.z * (1 - w));
  }
  static Float SchlickR0Eta(Float eta) { return Sqr((eta - 1) / (eta + 1)); }
  Spectrum DisneyFresnel(Float dielectricF, Float cosThetaI, bool isFrontSide, Float bsdfW, Float lum) const {
    bool isOutside = cosThetaI >= 0;
    Float rcpEta = Rcp(eta);
    Float etaIT = isOutside ? eta : rcpEta;
    Spectrum schlickF(0);
    if (hasMetallic) schlickF += SchlickFS(baseColor, cosThetaI, eta) * metallic;
    if (hasSpecTint) {
      Spectrum cTint = lum > 0 ? baseColor / lum : Spectrum(1);
      Spectrum F0 = cTint * SchlickR0Eta(etaIT);
      schlickF += SchlickFS(F0, cosThetaI, eta) * ((1 - metallic) * specTint);
    }
    Spectrum frontF = Spectrum((1 - metallic) * (1 - specTint) * dielectricF) + schlickF;
    return isFrontSide ? frontF : Spectrum(bsdfW * dielectricF);
  }
  static Float SmithGGX1(const V3& v, const V3& wh, Float alpha) {
    Float alpha2 = Sqr(alpha);
    Float cosTheta = std::abs(Frame::CosTheta(v));
    Float cosTheta2 = Sqr(cosTheta);
    Float tanTheta2 = (1 - cosTheta2) / cosTheta2;
    Float result = 2 * Rcp(1 + std::sqrt(1 + alpha2 * tanTheta2));
    if (Frame::CosTheta(v) == 1) result = 1;
    if (Dot(v, wh) * Frame::CosTheta(v) <= 0) result = 0;
    return result;
  }

  struct Probs {
    Float spec, trans, cc, diff;
  };
  Probs LobeProbs(bool isFrontSide, Float fr, Float brdfW, Float bsdfW) const {
    Float probSpecReflect = isFrontSide ? specularSample * (1 - bsdfW * (1 - fr)) : fr;
    Float probSpecTrans = hasSpecTrans ? (isFrontSide ? specularSample * bsdfW * (1 - fr) : (1 - fr)) : 0;
    Float probClearcoat = hasClearcoat ? (isFrontSide ? Float(0.25) * clearcoat * clearcoatSample : 0) : 0;
    Float probDiffuse = isFrontSide ? brdfW * diffuseReflectSample : 0;
    return {probSpecReflect, probSpecTrans, probClearcoat, probDiffuse};
  }

  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float lobeXi, const V2& dirXi) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi);
    BsdfSampleResult bsr{};
    if (cosThetaI == 0) return {bsr, Spectrum(0)};
    Float st = hasSpecTrans ? specTrans : 0;
    Float mt = hasMetallic ? metallic : 0;
    Float brdfW = (1 - mt) * (1 - st);
    Float bsdfW = hasSpecTrans ? (1 - mt) * st : 0;
    bool isFrontSide = cosThetaI > 0;
    Microfacet specDistr = SpecDist();
    V3 mSpec = specDistr.Sample(MulSign(si.Wi, cosThetaI), dirXi).first;
    DielectricF fd = FresnelDielectric(Dot(si.Wi, mSpec), eta);
    if (!isFrontSide && bsdfW <= 0) return {bsr, Spectrum(0)};
    Probs p = LobeProbs(isFrontSide, fd.r, brdfW, bsdfW);
    Float rcpTotal = Rcp(p.spec + p.trans + p.cc + p.diff);
    p.trans *= rcpTotal;
    p.cc *= rcpTotal;
    p.diff *= rcpTotal;
    Float cur = 0;
    bool sampleDiffuse = lobeXi < p.diff;
    cur += p.diff;
    bool sampleClearcoat = hasClearcoat && lobeXi >= cur && lobeXi < cur + p.cc;
    cur += p.cc;
    bool sampleSpecTrans = hasSpecTrans && lobeXi >= cur && lobeXi < cur + p.trans;
    cur += p.trans;
    bool sampleSpecReflect = lobeXi >= cur;
    bsr.Eta = 1;
    if (sampleSpecReflect) {
      V3 wo = Reflect(si.Wi, mSpec);
      bsr.Wo = wo;
      bsr.TypeMask = BSDF_REFLECTION | BSDF_GLOSSY;
      bool reflect = cosThetaI * Frame::CosTheta(wo) > 0;
      if (!(CheckMacMicNormal(mSpec, si.Wi, wo, cosThetaI, true) && reflect)) return {bsr, Spectrum(0)};
    }
    if (sampleSpecTrans) {
      V3 wo = Refract(si.Wi, mSpec, fd.cosThetaT, fd.etaTI);
      bsr.Wo = wo;
      bsr.TypeMask = BSDF_TRANSMISSION | BSDF_GLOSSY;
      bsr.Eta = fd.etaIT;
      bool refract = cosThetaI * Frame::CosTheta(wo) < 0;
      if (!(CheckMacMicNormal(mSpec, si.Wi, wo, cosThetaI, false) && refract)) return {bsr, Spectrum(0)};
    }
    if (hasClearcoat && sampleClearcoat) {
      GTR1 ccDist{Lerp(Float(0.1), Float(0.001), clearcoatGloss)};
      V3 mCC = ccDist.Sample(dirXi);
      V3 wo = Reflect(si.Wi, mCC);
      bsr.Wo = wo;
      bsr.TypeMask = BSDF_REFLECTION | BSDF_GLOSSY;
      bool reflect = cosThetaI * Frame::CosTheta(wo) > 0;
      if (!(CheckMacMicNormal(mCC, si.Wi, wo, cosThetaI, true) && reflect)) return {bsr, Spectrum(0)};
    }
    if (sampleDiffuse) {
      V3 wo = Warp::SquareToCosineHemisphere(dirXi);
      bsr.Wo = wo;
      bsr.TypeMask = BSDF_REFLECTION | BSDF_DIFFUSE;
      bool reflect = cosThetaI * Frame::CosTheta(wo) > 0;
      if (!reflect) return {bsr, Spectrum(0)};
    }
    bsr.Pdf = Pdf(si, bsr.Wo);
    if (bsr.Pdf <= 0) return {bsr, Spectrum(0)};
    return {bsr, Eval(si, bsr.Wo)};
  }

  Spectrum Eval(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi);
    if (cosThetaI == 0) return Spectrum(0);
    Float st = hasSpecTrans ? specTrans : 0;
    Float mt = hasMetallic ? metallic : 0;
    Float cc = hasClearcoat ? clearcoat : 0;
    Float sh = hasSheen ? sheen : 0;
    Float fl = hasFlatness ? flatness : 0;
    Float brdfW = (1 - mt) * (1 - st);
    Float bsdfW = (1 - mt) * st;
    Float cosThetaO = Frame::CosTheta(wo);
    bool reflect = cosThetaI * cosThetaO > 0;
    bool refract = cosThetaI * cosThetaO < 0;
    bool isFrontSide = cosThetaI > 0;
    Float invEta = Rcp(eta);
    Float etaPath = isFrontSide ? eta : invEta;
    Float invEtaPath = isFrontSide ? invEta : eta;
    Microfacet specDistr = SpecDist();
    V3 wh = Normalized(si.Wi + wo * (reflect ? Float(1) : etaPath));
    wh = MulSign(wh, Frame::CosTheta(wh));
    DielectricF fd = FresnelDielectric(Dot(si.Wi, wh), eta);
    bool canReflection = CheckMacMicNormal(wh, si.Wi, wo, cosThetaI, true);
    bool canRefraction = CheckMacMicNormal(wh, si.Wi, wo, cosThetaI, false);
    bool hasSpecReflectLobe = reflect && canReflection && fd.r > 0;
    bool hasClearcoatLobe = hasClearcoat && cc > 0 && reflect && canReflection && isFrontSide;
    bool hasSpecTransLobe = hasSpecTrans && bsdfW > 0 && refract && canRefraction && fd.r < 1;
    bool hasDiffuseLobe = brdfW > 0 && reflect && isFrontSide;
    bool hasSheenLobe = hasSheen && sh > 0 && reflect && (1 - mt) > 0 && isFrontSide;
    Float D = specDistr.D(wh);
    Float G = specDistr.G(si.Wi, wo, wh);
    Spectrum value(0);
    if (hasSpecReflectLobe) {
      Float lum = hasSpecTint ? Luminance(baseColor) : Float(1);
      Float specTintV = hasSpecTint ? specTint : 0;
      Disney tmp = *this;  // DisneyFresnel reads metallic / specTint as passed by the caller
      tmp.metallic = mt;
      tmp.specTint = specTintV;
      Spectrum F = tmp.DisneyFresnel(fd.r, Dot(si.Wi, wh), isFrontSide, bsdfW, lum);
      value += F * D * G / (Float(4) * std::abs(cosThetaI));
    }
    if (hasSpecTrans && hasSpecTransLobe) {
      Float scale = Sqr(invEtaPath);
      Float idh = Dot(si.Wi, wh), odh = Dot(wo, wh);
      Float w = bsdfW * std::abs((scale * (1 - fd.r) * D * G * etaPath * etaPath * idh * odh) /
                                 (cosThetaI * Sqr(idh + etaPath * odh)));
      value += Sqrt(baseColor) * w;
    }
    if (hasClearcoat && hasClearcoatLobe) {
      Float ccF = SchlickF(Float(0.04), Dot(si.Wi, wh), eta);
      GTR1 ccDistr{Lerp(Float(0.1), Float(0.001), clearcoatGloss)};
      Float ccD = ccDistr.Eval(wh);
      Float ccG = SmithGGX1(si.Wi, wh, 0.25f) * SmithGGX1(wo, wh, 0.25f);
      value += Spectrum((cc * Float(0.25)) * ccF * ccD * ccG * std::abs(cosThetaO));
    }
    if (hasDiffuseLobe) {
      Float Fo = SchlickWeight(std::abs(cosThetaO));
      Float Fi = SchlickWeight(std::abs(cosThetaI));
      Float fsDiff = (1 - Float(0.5) * Fi) * (1 - Float(0.5) * Fo);
      Float cosThetaD = Dot(wh, wo);
      Float Rr = 2 * roughness * Sqr(cosThetaD);
      Float fsRetro = Rr * (Fo + Fi + Fo * Fi * (Rr - 1));
      if (hasFlatness) {
        Float Fss90 = Rr / 2;
        Float Fss = Lerp(1, Fss90, Fo) * Lerp(1, Fss90, Fi);
        Float fsSSS = Float(1.25) * (Fss * (1 / (std::abs(cosThetaO) + std::abs(cosThetaI)) - Float(0.5)) + Float(0.5));
        value += baseColor * (brdfW * std::abs(cosThetaO)) * (1 / PI) * Lerp(fsDiff + fsRetro, fsSSS, fl);
      } else {
        value += baseColor * (brdfW * std::abs(cosThetaO)) * (1 / PI) * (fsDiff + fsRetro);
      }
      if (hasSheen && hasSheenLobe) {
        Float Fd = SchlickWeight(std::abs(cosThetaD));
        if (hasSheenTint) {
          Float lum = Luminance(baseColor);
          Spectrum cTint = lum > 0 ? baseColor / lum : Spectrum(1);
          Spectrum cSheen = Spectrum(1) + (cTint - Spectrum(1)) * sheenTint;
          value += cSheen * (sh * (1 - mt) * Fd * std::abs(cosThetaO));
        } else {
          value += Spectrum(sh * (1 - mt) * Fd * std::abs(cosThetaO));
        }
      }
    }
    return value;
  }

  Float Pdf(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi);
    if (cosThetaI == 0) return 0;
    Float st = hasSpecTrans ? specTrans : 0;
    Float mt = hasMetallic ? metallic : 0;
    Float brdfW = (1 - mt) * (1 - st);
    Float bsdfW = (1 - mt) * st;
    bool isFrontSide = cosThetaI > 0;
    Float etaPath = isFrontSide ? eta : Rcp(eta);
    Float cosThetaO = Frame::CosTheta(wo);
    bool reflect = cosThetaI * cosThetaO > 0;
    bool refract = cosThetaI * cosThetaO < 0;
    V3 wh = Normalized(si.Wi + wo * (reflect ? Float(1) : etaPath));
    wh = MulSign(wh, Frame::CosTheta(wh));
    Microfacet specDistr = SpecDist();
    DielectricF fd = FresnelDielectric(Dot(si.Wi, wh), eta);
    Probs p = LobeProbs(isFrontSide, fd.r, brdfW, bsdfW);
    Float rcpTotal = Rcp(p.spec + p.trans + p.cc + p.diff);
    p.spec *= rcpTotal;
    p.trans *= rcpTotal;
    p.cc *= rcpTotal;
    p.diff *= rcpTotal;
    Float dwhdwoAbs;
    if (hasSpecTrans) {
      Float wiDotH = Dot(si.Wi, wh), woDotH = Dot(wo, wh);
      dwhdwoAbs = std::abs(reflect ? Rcp(Float(4) * woDotH) : (Sqr(etaPath) * woDotH) / Sqr(wiDotH + etaPath * woDotH));
    } else {
      dwhdwoAbs = std::abs(Rcp(Float(4) * Dot(wo, wh)));
    }
    Float pdf = 0;
    bool canReflection = CheckMacMicNormal(wh, si.Wi, wo, cosThetaI, true) && reflect;
    if (canReflection) pdf += p.spec * specDistr.Pdf(MulSign(si.Wi, cosThetaI), wh) * dwhdwoAbs;
    if (reflect) pdf += p.diff * Warp::SquareToCosineHemispherePdf(wo);
    if (hasSpecTrans) {
      bool canRefr = CheckMacMicNormal(wh, si.Wi, wo, cosThetaI, false) && refract;
      if (canRefr) pdf += p.trans * specDistr.Pdf(MulSign(si.Wi, cosThetaI), wh) * dwhdwoAbs;
    }
    if (hasClearcoat) {
      GTR1 ccDistr{Lerp(Float(0.1), Float(0.001), clearcoatGloss)};
      if (canReflection) pdf += p.cc * ccDistr.Pdf(wh) * dwhdwoAbs;
    }
    return pdf;
  }
};

// ---------------------------------------------------------------- plastic (bsdf/plastic.cpp:18-132)
// Fresnel::DiffuseReflectance (fresnel.cpp:79-84): the polynomial fit of the
// hemispherical Fresnel average, Horner-evaluated with fused multiply-adds
// as math_base.h:90-105 does.
static Float DiffuseReflectance(Float eta) {
  Float invEta = Rcp(eta);
  if (eta < 1) return Fmadd(Float(0.0636f), invEta, Fmadd(eta, Fmadd(eta, Float(-1.4399f), Float(0.7099f)), Float(0.6681f)));
  const Float c[6] = {0.919317f, -3.4793f, 6.75335f, -7.80989f, 4.98554f, -1.36881f};
  Float acc = c[5];
  for (int i = 4; i >= 0; i--) acc = Fmadd(invEta, acc, c[i]);
  return acc;
}

// Shared by plastic and rough_plastic: sampling weight from the luminance of
// the (constant) diffuse and specular colours, and the diffuse term under
// the dielectric coating.
struct PlasticBase : Bsdf {
  Float eta = 1.49f / 1.000277f, fdrInt = 0, invEta2 = 1, specularSampleWeight = 0.5f;
  Spectrum diffuse, specular;
  bool nonlinear = false;
  void Init() {
    fdrInt = DiffuseReflectance(Float(1) / eta);
    invEta2 = Float(1) / (eta * eta);
    double dm = Luminance(diffuse), sm = Luminance(specular);
    specularSampleWeight = Float(sm / (dm + sm));
  }
  Spectrum DiffuseTerm(Float fi, Float fo, Float cosThetaO) const {
    Spectrum num = diffuse * (1 / PI) * invEta2 * (1 - fi) * (1 - fo) * cosThetaO;
    if (nonlinear) {
      return Spectrum(num.x * (1 / (1 - diffuse.x * fdrInt)), num.y * (1 / (1 - diffuse.y * fdrInt)),
                      num.z * (1 / (1 - diffuse.z * fdrInt)));
    }
    return num / (1 - fdrInt);
  }
};

struct Plastic final : PlasticBase {
  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float lobeXi, const V2& dirXi) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi);
    if (cosThetaI <= 0) return {{}, Spectrum(0)};
    Float fi = FresnelDielectric(cosThetaI, eta).r;
    Float probSpec = fi * specularSampleWeight;
    Float probDiff = (1 - fi) * (1 - specularSampleWeight);
    probSpec = probSpec / (probSpec + probDiff);
    probDiff = 1 - probSpec;
    BsdfSampleResult bsr{};
    bsr.Eta = 1;
    if (lobeXi < probSpec) {
      bsr.Wo = V3(-si.Wi.x, -si.Wi.y, si.Wi.z);
      bsr.Pdf = probSpec;
      bsr.TypeMask = BSDF_REFLECTION | BSDF_DELTA;
      return {bsr, specular * fi};
    }
    bsr.Wo = Warp::SquareToCosineHemisphere(dirXi);
    bsr.Pdf = Warp::SquareToCosineHemispherePdf(bsr.Wo) * probDiff;
    bsr.TypeMask = BSDF_REFLECTION | BSDF_DIFFUSE;
    Float cosThetaO = Frame::CosTheta(bsr.Wo);
    Float fo = FresnelDielectric(cosThetaO, eta).r;
    return {bsr, DiffuseTerm(fi, fo, cosThetaO)};
  }
  Spectrum Eval(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi), cosThetaO = Frame::CosTheta(wo);
    if (cosThetaI <= 0 || cosThetaO <= 0) return Spectrum(0);
    Float fi = FresnelDielectric(cosThetaI, eta).r;
    Float fo = FresnelDielectric(cosThetaO, eta).r;
    return DiffuseTerm(fi, fo, cosThetaO);
  }
  Float Pdf(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi), cosThetaO = Frame::CosTheta(wo);
    if (cosThetaI <= 0 || cosThetaO <= 0) return 0;
    Float fi = FresnelDielectric(cosThetaI, eta).r;
    Float probSpec = fi * specularSampleWeight;
    Float probDiff = (1 - fi) * (1 - specularSampleWeight);
    probDiff = probDiff / (probDiff + probSpec);
    return Warp::SquareToCosineHemispherePdf(wo) * probDiff;
  }
};

// ---------------------------------------------------------------- rough_plastic (bsdf/rough_plastic.cpp:18-250)
struct RoughPlastic final : PlasticBase {
  Microfacet dist;
  Float ProbSpec(Float fi) const {
    Float probSpec = fi * specularSampleWeight;
    Float probDiff = (1 - fi) * (1 - specularSampleWeight);
    return probSpec / (probSpec + probDiff);
  }
  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float lobeXi, const V2& dirXi) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi);
    if (cosThetaI <= 0) return {{}, Spectrum(0)};
    Float probSpec = ProbSpec(FresnelDielectric(cosThetaI, eta).r);
    BsdfSampleResult bsr{};
    bsr.Eta = 1;
    if (lobeXi < probSpec) {
      auto [wh, pdf] = dist.Sample(si.Wi, dirXi);
      if (pdf <= 0) return {bsr, Spectrum(0)};
      bsr.Wo = Reflect(si.Wi, wh);
      bsr.TypeMask = BSDF_GLOSSY | BSDF_REFLECTION;
    } else {
      bsr.Wo = Warp::SquareToCosineHemisphere(dirXi);
      bsr.TypeMask = BSDF_DIFFUSE | BSDF_REFLECTION;
    }
    bsr.Pdf = Pdf(si, bsr.Wo);
    if (bsr.Pdf <= 0) return {bsr, Spectrum(0)};
    return {bsr, Eval(si, bsr.Wo)};
  }
  Spectrum Eval(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi), cosThetaO = Frame::CosTheta(wo);
    if (cosThetaI <= 0 || cosThetaO <= 0) return Spectrum(0);
    V3 wh = Normalized(wo + si.Wi);
    Float F = FresnelDielectric(Dot(si.Wi, wh), eta).r;
    Spectrum spec = specular * dist.D(wh) * dist.G(si.Wi, wo, wh) * F / (4 * cosThetaI);
    Float fo = FresnelDielectric(Dot(wo, wh), eta).r;
    return spec + DiffuseTerm(F, fo, cosThetaO);
  }
  Float Pdf(const SurfaceInteraction& si, const V3& wo) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi), cosThetaO = Frame::CosTheta(wo);
    if (cosThetaI <= 0 || cosThetaO <= 0) return 0;
    Float probSpec = ProbSpec(FresnelDielectric(cosThetaI, eta).r);
    Float probDiff = 1 - probSpec;
    V3 wh = Normalized(wo + si.Wi);
    Float specPdf = dist.Pdf(si.Wi, wh) / (4 * Dot(wo, wh)) * probSpec;
    Float diffPdf = Warp::SquareToCosineHemispherePdf(wo) * probDiff;
    return specPdf + diffPdf;
  }
};

// ---------------------------------------------------------------- perfect_mirror (bsdf/perfect_mirror.cpp:16-62)
struct PerfectMirror final : Bsdf {
  Spectrum reflectance, eta, k;
  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float, const V2&) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi);
    if (cosThetaI <= 0) return {{}, Spectrum(0)};
    BsdfSampleResult bsr{};
    bsr.Wo = V3(-si.Wi.x, -si.Wi.y, si.Wi.z);
    bsr.Eta = 1;
    bsr.TypeMask = flags;
    bsr.Pdf = 1;
    return {bsr, Mul(reflectance, FresnelConductor(cosThetaI, eta, k))};
  }
  Spectrum Eval(const SurfaceInteraction&, const V3&) const override { return Spectrum(0); }
  Float Pdf(const SurfaceInteraction&, const V3&) const override { return 0; }
};

// ---------------------------------------------------------------- perfect_glass (bsdf/perfect_glass.cpp:14-83)
struct PerfectGlass final : Bsdf {
  Spectrum reflectance, transmittance;
  Float eta = 1.5046f / 1.000277f;
  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float lobeXi, const V2&) const override {
    Float cosThetaI = Frame::CosTheta(si.Wi);
    DielectricF fr = FresnelDielectric(cosThetaI, eta);
    Float ri = fr.r, ti = 1 - ri;
    BsdfSampleResult bsr{};
    if (lobeXi <= ri) {
      bsr.Wo = V3(-si.Wi.x, -si.Wi.y, si.Wi.z);
      bsr.Eta = 1;
      bsr.Pdf = ri;
      bsr.TypeMask = BSDF_DELTA | BSDF_REFLECTION;
      return {bsr, reflectance * ri};
    }
    bsr.Wo = V3(-fr.etaTI * si.Wi.x, -fr.etaTI * si.Wi.y, fr.cosThetaT);
    bsr.Eta = fr.etaIT;
    bsr.Pdf = ti;
    bsr.TypeMask = BSDF_DELTA | BSDF_TRANSMISSION;
    Spectrum f = transmittance * ti;
    return {bsr, f * Sqr(fr.etaTI)};
  }
  Spectrum Eval(const SurfaceInteraction&, const V3&) const override { return Spectrum(0); }
  Float Pdf(const SurfaceInteraction&, const V3&) const override { return 0; }
};

// ---------------------------------------------------------------- two_side (bsdf/two_side.cpp:60-137, "two_side" form)
struct TwoSide final : Bsdf {
  std::unique_ptr<Bsdf> nested;
  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float lobeXi, const V2& dirXi) const override {
    Float c = Frame::CosTheta(si.Wi);
    if (c > 0) return nested->Sample(si, lobeXi, dirXi);
    if (c < 0) {
      SurfaceInteraction t = si;
      t.Wi.z = -t.Wi.z;
      auto [bsr, f] = nested->Sample(t, lobeXi, dirXi);
      bsr.Wo.z = -bsr.Wo.z;
      return {bsr, f};
    }
    BsdfSampleResult bsr{};
    bsr.Pdf = 0;
    return {bsr, Spectrum(0)};
  }
  Spectrum Eval(const SurfaceInteraction& si, const V3& wo) const override {
    Float c = Frame::CosTheta(si.Wi);
    if (c > 0) return nested->Eval(si, wo);
    if (c < 0) {
      SurfaceInteraction t = si;
      t.Wi.z = -t.Wi.z;
      return nested->Eval(t, V3(wo.x, wo.y, -wo.z));
    }
    return Spectrum(0);
  }
  Float Pdf(const SurfaceInteraction& si, const V3& wo) const override {
    Float c = Frame::CosTheta(si.Wi);
    if (c > 0) return nested->Pdf(si, wo);
    if (c < 0) {
      SurfaceInteraction t = si;
      t.Wi.z = -t.Wi.z;
      return nested->Pdf(t, V3(wo.x, wo.y, -wo.z));
    }
    return 0;
  }
};

// ---------------------------------------------------------------- mask (bsdf/mask.cpp:8-62)
// The pass-through sample keeps the reference's bookkeeping: type Null (not
// delta), pdf 1 - opacity, weight 1.
struct Mask final : Bsdf {
  std::unique_ptr<Bsdf> nested;
  Float opacity = 0.5f;
  std::pair<BsdfSampleResult, Spectrum> Sample(const SurfaceInteraction& si, Float lobeXi, const V2& dirXi) const override {
    if (lobeXi < opacity) {
      lobeXi /= opacity;
      auto [bsr, f] = nested->Sample(si, lobeXi, dirXi);
      bsr.Pdf *= opacity;
      return {bsr, f * opacity};
    }
    BsdfSampleResult bsr{};
    bsr.Pdf = 1 - opacity;
    bsr.Eta = 1;
    bsr.Wo = -si.Wi;
    bsr.TypeMask = 0;
    return {bsr, Spectrum(1)};
  }
  Spectrum Eval(const SurfaceInteraction& si, const V3& wo) const override { return nested->Eval(si, wo) * opacity; }
  Float Pdf(const SurfaceInteraction& si, const V3& wo) const override { return nested->Pdf(si, wo) * opacity; }
};

float RefErfInv(float a) { return ErfInv(a); }

std::unique_ptr<Bsdf> MakeBsdf(const rad_bsdf_desc& d, bool quirks) { return MakeBsdf(&d, 1, 0, quirks, 0); }

std::unique_ptr<Bsdf> MakeBsdf(const rad_bsdf_desc* all, uint32_t count, uint32_t index, bool quirks, int depth,
                               const Float* specWeights) {
  if (index >= count || depth > 2) throw std::runtime_error("bsdf: nested index out of range");
  const rad_bsdf_desc& d = all[index];
  auto spec = [](const float* v) { return Spectrum(v[0], v[1], v[2]); };
  Microfacet md;
  md.beckmann = d.distribution == RAD_MICROFACET_BECKMANN;
  md.alphaX = d.alpha_u;
  md.alphaY = d.alpha_v;
  md.sampleVisible = d.sample_visible != 0;
  md.quirks = quirks;
  std::unique_ptr<Bsdf> out;
  switch (d.type) {
    case RAD_BSDF_DIFFUSE: {
      auto b = std::make_unique<Diffuse>();
      b->flags = BSDF_DIFFUSE | BSDF_REFLECTION;
      b->reflectance = spec(d.reflectance);
      out = std::move(b);
      break;
    }
    case RAD_BSDF_ROUGH_METAL: {
      auto b = std::make_unique<RoughMetal>();
      b->flags = BSDF_GLOSSY | BSDF_REFLECTION;
      b->reflectance = spec(d.reflectance);
      b->eta = spec(d.eta);
      b->k = spec(d.k);
      b->dist = md;
      out = std::move(b);
      break;
    }
    case RAD_BSDF_ROUGH_GLASS: {
      auto b = std::make_unique<RoughGlass>();
      b->flags = BSDF_GLOSSY | BSDF_REFLECTION | BSDF_TRANSMISSION;
      b->reflectance = spec(d.reflectance);
      b->transmittance = spec(d.transmittance);
      b->eta = d.int_ior / d.ext_ior;
      b->invEta = d.ext_ior / d.int_ior;
      b->dist = md;
      out = std::move(b);
      break;
    }
    case RAD_BSDF_DISNEY: {
      auto b = std::make_unique<Disney>();
      uint32_t f = d.disney_flags;
      b->baseColor = spec(d.base_color);
      b->roughness = d.roughness;
      b->hasAnisotropic = f & RAD_DISNEY_HAS_ANISOTROPIC;
      b->anisotropic = d.anisotropic;
      b->hasSpecTrans = f & RAD_DISNEY_HAS_SPEC_TRANS;
      b->specTrans = d.spec_trans;
      b->hasSheen = f & RAD_DISNEY_HAS_SHEEN;
      b->sheen = d.sheen;
      b->hasSheenTint = f & RAD_DISNEY_HAS_SHEEN_TINT;
      b->sheenTint = d.sheen_tint;
      b->hasFlatness = f & RAD_DISNEY_HAS_FLATNESS;
      b->flatness = d.flatness;
      b->hasSpecTint = f & RAD_DISNEY_HAS_SPEC_TINT;
      b->specTint = d.spec_tint;
      b->hasMetallic = f & RAD_DISNEY_HAS_METALLIC;
      b->metallic = d.metallic;
      b->hasClearcoat = f & RAD_DISNEY_HAS_CLEARCOAT;
      b->clearcoat = d.clearcoat;
      b->clearcoatGloss = d.clearcoat_gloss;
      b->specularSample = d.specular_sampling_rate;
      b->clearcoatSample = d.clearcoat_sampling_rate;
      b->diffuseReflectSample = d.diffuse_sampling_rate;
      if (f & RAD_DISNEY_HAS_ETA) {  // disney.cpp:86-101
        b->eta = d.disney_eta;
        if (b->hasSpecTrans && b->eta == 1) b->eta = Float(1.001);
      } else {
        Float specular = d.specular;
        if (b->hasSpecTrans && specular == 0) specular = Float(0.001);
        b->eta = Float(2) * Rcp(Float(1) - std::sqrt(Float(0.08) * specular)) - Float(1);
      }
      b->flags = BSDF_DIFFUSE | BSDF_REFLECTION;
      if (b->hasClearcoat) b->flags |= BSDF_GLOSSY;
      if (b->hasSpecTrans) b->flags |= BSDF_TRANSMISSION;
      out = std::move(b);
      break;
    }
    case RAD_BSDF_PLASTIC:
    case RAD_BSDF_ROUGH_PLASTIC: {
      std::unique_ptr<PlasticBase> b;
      if (d.type == RAD_BSDF_PLASTIC) {
        b = std::make_unique<Plastic>();
      } else {
        auto r = std::make_unique<RoughPlastic>();
        r->dist = md;
        r->dist.alphaY = md.alphaX;  // isotropic "alpha"
        b = std::move(r);
      }
      b->flags = BSDF_DIFFUSE | BSDF_REFLECTION | BSDF_DELTA;  // both declare Delta (plastic.cpp:22, rough_plastic.cpp:22)
      b->eta = d.int_ior / d.ext_ior;
      b->diffuse = spec(d.diffuse);
      b->specular = spec(d.specular_reflectance);
      b->nonlinear = d.nonlinear != 0;
      b->Init();
      if (specWeights && !std::isnan(specWeights[index])) b->specularSampleWeight = specWeights[index];
      out = std::move(b);
      break;
    }
    case RAD_BSDF_PERFECT_MIRROR: {
      auto b = std::make_unique<PerfectMirror>();
      b->flags = BSDF_DELTA | BSDF_REFLECTION;
      b->reflectance = spec(d.reflectance);
      b->eta = spec(d.eta);
      b->k = spec(d.k);
      out = std::move(b);
      break;
    }
    case RAD_BSDF_PERFECT_GLASS: {
      auto b = std::make_unique<PerfectGlass>();
      b->flags = BSDF_DELTA | BSDF_REFLECTION | BSDF_TRANSMISSION;
      b->reflectance = spec(d.reflectance);
      b->transmittance = spec(d.transmittance);
      b->eta = d.int_ior / d.ext_ior;
      out = std::move(b);
      break;
    }
    case RAD_BSDF_TWO_SIDE: {
      auto b = std::make_unique<TwoSide>();
      if (d.nested < 0) throw std::runtime_error("two_side: no nested bsdf");
      b->nested = MakeBsdf(all, count, uint32_t(d.nested), quirks, depth + 1, specWeights);
      b->flags = b->nested->flags;
      out = std::move(b);
      break;
    }
    case RAD_BSDF_MASK: {
      auto b = std::make_unique<Mask>();
      if (d.nested < 0) throw std::runtime_error("mask: no nested bsdf");
      b->nested = MakeBsdf(all, count, uint32_t(d.nested), quirks, depth + 1, specWeights);
      b->opacity = std::clamp(d.opacity, Float(0), Float(1));
      b->flags = b->nested->flags;
      out = std::move(b);
      break;
    }
    default:
      throw std::runtime_error("unsupported bsdf type");
  }
  out->quirks = quirks;
  return out;
}

}  // namespace ref
