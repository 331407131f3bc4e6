// Device BSDFs on the path: diffuse, rough_metal, rough_glass, disney,
// plastic, rough_plastic, perfect_mirror, perfect_glass and the two_side /
// mask wrappers, with the GGX / Beckmann microfacet distributions and
// Fresnel terms.  Directions are in
// the shading frame (z = shading normal); `wi` points away from the surface
// toward the previous vertex (SurfaceInteraction::Wi).
#pragma once

#include "scene_dev.h"

namespace rad {

struct BsdfSample {
  f3 wo;
  float pdf;
  float eta;
  uint32_t lobes;
};

// ---------------------------------------------------------------- Fresnel (fresnel.cpp:10-77)
RAD_HD f3 reflect_wh(f3 wi, f3 wh) { return fmadd(wh, splat(2.0f * dot(wi, wh)), -wi); }
RAD_HD f3 refract_wh(f3 wi, f3 wh, float cosThetaT, float etaTI) {
  return fmadd(wh, splat(fmaf(dot(wi, wh), etaTI, cosThetaT)), -(wi * etaTI));
}
struct Dielectric {
  float r, cos_t, eta_it, eta_ti;
};
RAD_HD Dielectric fresnel_dielectric(float cosThetaI, float eta) {
  bool outside = cosThetaI >= 0.f;
  float rcpEta = 1.0f / eta;
  float etaIT = outside ? eta : rcpEta;
  float etaTI = outside ? rcpEta : eta;
  float sqrCosT = fmaf(-fmaf(-cosThetaI, cosThetaI, 1.0f), etaTI * etaTI, 1.0f);
  float aci = fabsf(cosThetaI);
  float act = safe_sqrt(sqrCosT);
  bool matched = eta == 1.0f;
  float aS = fmaf(-etaIT, act, aci) / fmaf(etaIT, act, aci);
  float aP = fmaf(-etaIT, aci, act) / fmaf(etaIT, aci, act);
  float r = 0.5f * (sqr(aS) + sqr(aP));
  if (matched || aci == 0.f) r = matched ? 0.f : 1.f;
  return {r, cosThetaI >= 0.f ? -act : act, etaIT, etaTI};
}
RAD_HD float conductor1(float cosThetaI, float etaR, float etaI) {
  float c2 = cosThetaI * cosThetaI, s2 = 1.0f - c2, s4 = s2 * s2;
  float temp1 = etaR * etaR - etaI * etaI - s2;
  float a2pb2 = sqrtf(fmaxf(temp1 * temp1 + etaI * etaI * etaR * etaR * 4.0f, 0.f));
  float a = sqrtf(fmaxf((a2pb2 + temp1) * 0.5f, 0.f));
  float term1 = a2pb2 + c2, term2 = a * 2.0f * cosThetaI;
  float rs = (term1 - term2) * (1.0f / (term1 + term2));
  float term3 = a2pb2 * c2 + s4, term4 = term2 * s2;
  float rp = rs * (term3 - term4) * (1.0f / (term3 + term4));
  return (rs + rp) * 0.5f;
}
RAD_HD f3 fresnel_conductor(float cosThetaI, f3 eta, f3 k) {
  return mk(conductor1(cosThetaI, eta.x, k.x), conductor1(cosThetaI, eta.y, k.y), conductor1(cosThetaI, eta.z, k.z));
}

// ---------------------------------------------------------------- microfacet (microfacet.cpp:12-277)
// erf^-1 for the Beckmann visible-normal sampler.  Like the reference's
// ErfInv (math_ext.h:114-140) the argument enters through 1 - a^2 clamped at
// FLT_MIN, so |a| >= 1 saturates instead of diverging.  Solved in double:
// with c = (1 - a^2) / (1 + sqrt(a^2)) = erfc(|y|), Newton on
// log(erfc(y)) = log(c), which is concave and decreasing, so the iteration
// converges monotonically after at most one overshoot.
// Kept out of line: only Beckmann visible-normal sampling calls it, and
// inlined, its double-precision erf / erfc / log / exp would hold registers
// in every shading kernel that can sample a microfacet (rough metal
// 232 -> 188 VGPRs, rough glass 238 -> 192, guided rough metal 256 -> 216).
__device__ __noinline__ float erfinv_newton(float a) {
  if (fabsf(a) < 0.5f) {  // well-conditioned core: Newton on erf(y) = |a|
    double A = fabs(double(a)), y = A * 0.886226925452758;
    for (int i = 0; i < 30; i++) {
      double step = (erf(y) - A) / (1.1283791670955126 * exp(-y * y));
      y -= step;
      if (fabs(step) <= 1e-16 * fmax(1e-300, y)) break;
    }
    return copysignf(float(y), a);
  }
  float q = fmaxf(fmaf(a, -a, 1.0f), 1.17549435e-38f);
  double Q = q;
  double c = Q / (1.0 + sqrt(1.0 - Q));
  if (c >= 1.0) return 0.0f * a;
  double L = log(c);
  double y = c > 0.5 ? (1.0 - c) * 0.886226925452758 : sqrt(fmax(-L - 0.5 * log(-L), 1e-6));
  for (int i = 0; i < 60; i++) {
    double e = erfc(y);
    double f = log(e) - L;
    double fp = -1.1283791670955126 * exp(-y * y) / e;
    double step = f / fp;
    y -= step;
    if (y < 0.0) y = 0.0;
    if (fabs(step) <= 1e-15 * fmax(1.0, y)) break;
  }
  return copysignf(float(y), a);
}

struct Microfacet {
  float ax, ay;
  bool beckmann;
  bool visible;
  bool quirks;

  __device__ float D(f3 wh) const {
    float axy = ax * ay;
    float ct = wh.z;
    if (beckmann) {
      float c2 = sqr(ct);
      float r = expf(-(sqr(wh.x / ax) + sqr(wh.y / ay)) / c2) / (kPi * axy * sqr(c2));
      return r * ct > 1e-20f ? r : 0.f;
    }
    float r = 1.0f / (kPi * axy * sqr(sqr(wh.x / ax) + sqr(wh.y / ay) + sqr(wh.z)));
    return r * ct > 1e-20f ? r : 0.f;
  }
  __device__ float G1(f3 v, f3 wh) const {
    float xy = sqr(ax * v.x) + sqr(ay * v.y);
    float t2 = xy / sqr(v.z);
    float r;
    if (beckmann) {
      float a = 1.0f / sqrtf(t2);
      float a2 = a * a;
      r = a >= 1.6f ? 1.f : (3.535f * a + 2.181f * a2) / (1.f + 2.276f * a + 2.577f * a2);
    } else {
      r = 2.0f / (1.0f + sqrtf(1.0f + t2));
    }
    if (xy == 0.f) r = 1.f;
    if (dot(v, wh) * v.z <= 0.f) r = 0.f;
    return r;
  }
  __device__ float G(f3 wi, f3 wo, f3 wh) const { return G1(wi, wh) * G1(wo, wh); }
  __device__ float Pdf(f3 wi, f3 wh) const {
    float r = D(wh);
    if (visible) {
      r *= G1(wi, wh) * fabsf(dot(wi, wh)) / wi.z;
    } else {
      r *= wh.z;
    }
    return r;
  }
  __device__ static void visible_ggx(float cosThetaI, float sx, float sy, float& ox, float& oy) {  // :202-236
    if (cosThetaI > 0.9999f) {
      float r = sqrtf(sx / (1.f - sx));
      float phi = 6.28318530718f * sy;
      ox = r * cosf(phi);
      oy = r * sinf(phi);
      return;
    }
    float sinTheta = sqrtf(fmaxf(0.f, 1.f - sqr(cosThetaI)));
    float tanTheta = sinTheta / cosThetaI;
    float a = 1.f / tanTheta;
    float G1v = 2.f / (1.f + sqrtf(1.f + 1.f / sqr(a)));
    float A = 2.f * sx / G1v - 1.f;
    float tmp = 1.f / fmaf(A, A, -1.f);
    if (tmp > 1e10f) tmp = 1e10f;
    float B = tanTheta;
    float Dd = sqrtf(fmaxf(sqr(B) * sqr(tmp) - (sqr(A) - sqr(B)) * tmp, 0.f));
    float s1 = fmaf(B, tmp, -Dd), s2 = fmaf(B, tmp, Dd);
    float slopeX = (A < 0.f || s2 > 1.f / tanTheta) ? s1 : s2;
    float S;
    if (sy > 0.5f) {
      S = 1.f;
      sy = 2.f * (sy - 0.5f);
    } else {
      S = -1.f;
      sy = 2.f * (0.5f - sy);
    }
    float z = (sy * (sy * (sy * 0.27385f - 0.73369f) + 0.46341f)) / (sy * (sy * (sy * 0.093073f + 0.309420f) - 1.f) + 0.597999f);
    ox = slopeX;
    oy = S * z * sqrtf(1.f + slopeX * slopeX);
  }
  __device__ static void visible_beckmann(float cosThetaI, float sx, float sy, float& ox, float& oy) {  // :82-97
    float tanThetaI = safe_sqrt(fmaf(-cosThetaI, cosThetaI, 1.f)) / cosThetaI;
    float cot = 1.f / tanThetaI;
    float maxval = erff(cot);
    sx = fmaxf(fminf(sx, 1.f - 1e-6f), 1e-6f);
    sy = fmaxf(fminf(sy, 1.f - 1e-6f), 1e-6f);
    float x = maxval - (maxval + 1.f) * erff(sqrtf(-logf(sx)));
    sx *= 1.f + maxval + (1.f / sqrtf(kPi)) * tanThetaI * expf(-sqr(cot));
    for (int i = 0; i < 3; i++) {
      float slope = erfinv_newton(x);
      float value = 1.f + x + (1.f / sqrtf(kPi)) * tanThetaI * expf(-sqr(slope)) - sx;
      float deriv = 1.f - slope * tanThetaI;
      x -= value / deriv;
    }
    ox = erfinv_newton(x);
    oy = erfinv_newton(fmaf(2.f, sy, -1.f));
  }
  // Returns the microfacet normal and its pdf (MicrofacetDistribution::Sample).
  __device__ f3 Sample(f3 wi, float u1, float u2, float& pdf) const {
    if (visible) {
      f3 wp = normalize(mk(ax * wi.x, ay * wi.y, wi.z));
      // Frame::SinCosPhi (frame.cpp:48-56).  The reference clamps with
      // cwiseMin(-1).cwiseMax(1), which yields (1, 1) off the pole; kept as-is
      // unless quirks are switched off.
      float s2 = fmaf(wp.x, wp.x, sqr(wp.y));
      float inv = 1.f / sqrtf(s2);
      float cx = wp.x * inv, cy = wp.y * inv;
      if (fabsf(s2) <= 4.f * kFloat32Eps) {
        cx = 1.f;
        cy = 0.f;
      } else if (quirks) {
        cx = fmaxf(fminf(cx, -1.f), 1.f);
        cy = fmaxf(fminf(cy, -1.f), 1.f);
      } else {
        cx = fminf(fmaxf(cx, -1.f), 1.f);
        cy = fminf(fmaxf(cy, -1.f), 1.f);
      }
      float sinPhi = cy, cosPhi = cx;
      float sx, sy;
      if (beckmann)
        visible_beckmann(wp.z, u1, u2, sx, sy);
      else
        visible_ggx(wp.z, u1, u2, sx, sy);
      float tx = fmaf(cosPhi, sx, -(sinPhi * sy)) * ax;
      float ty = fmaf(sinPhi, sx, cosPhi * sy) * ay;
      f3 m = normalize(mk(-tx, -ty, 1.f));
      pdf = D(m) * G1(wi, m) * fabsf(dot(wi, m)) / wi.z;
      return m;
    }
    float sinPhi, cosPhi, alpha2;
    if (ax == ay) {
      sinPhi = sinf((2.f * kPi) * u2);
      cosPhi = cosf((2.f * kPi) * u2);
      alpha2 = ax * ay;
    } else {
      float ratio = ay / ax;
      float tmp = ratio * tanf((2.f * kPi) * u2);
      cosPhi = 1.f / sqrtf(fmaf(tmp, tmp, 1.f));
      cosPhi = mulsign(cosPhi, fabsf(u2 - 0.5f) - 0.25f);
      sinPhi = cosPhi * tmp;
      alpha2 = 1.f / (sqr(cosPhi / ax) + sqr(sinPhi / ay));
    }
    if (beckmann) {
      float ct = 1.f / sqrtf(fmaf(-alpha2, logf(1.f - u1), 1.f));
      float c2 = sqr(ct);
      float c3 = fmaxf(c2 * ct, 1e-20f);
      pdf = (1.f - u1) / (kPi * ax * ay * c3);
      float st = sqrtf(1.f - c2);
      return mk(cosPhi * st, sinPhi * st, ct);
    }
    float tanM2 = alpha2 * u1 / (1.f - u1);
    float ct = 1.f / sqrtf(1.f + tanM2);
    float c2 = sqr(ct);
    float temp = 1.f + tanM2 / alpha2;
    float c3 = fmaxf(c2 * ct, 1e-20f);
    pdf = 1.f / (kPi * ax * ay * c3 * sqr(temp));
    float st = sqrtf(1.f - c2);
    return mk(cosPhi * st, sinPhi * st, ct);
  }
};

__device__ inline Microfacet mf_of(const BsdfRec& b, bool quirks) {
  return Microfacet{b.alpha_u, b.alpha_v, b.beckmann != 0, b.sample_visible != 0, quirks};
}

// ---------------------------------------------------------------- diffuse (diffuse.cpp:40-99)
__device__ inline f3 diffuse_eval(const BsdfRec& b, f3 wi, f3 wo, float& pdf) {
  if (wi.z <= 0.f || wo.z <= 0.f) {
    pdf = 0.f;
    return splat(0.f);
  }
  pdf = wo.z * kInvPi;
  return b.reflectance * kInvPi * wo.z;
}
__device__ inline f3 diffuse_sample(const BsdfRec& b, f3 wi, float u1, float u2, BsdfSample& s) {
  if (wi.z <= 0.f) {
    s.pdf = 0.f;
    return splat(0.f);
  }
  float r = sqrtf(u1), ang = 2.f * kPi * u2;  // SquareToCosineHemisphere via the disk (warp.cpp:33-50)
  float x = r * cosf(ang), y = r * sinf(ang);
  s.wo = mk(x, y, safe_sqrt(1.f - sqr(x) - sqr(y)));
  s.pdf = s.wo.z * kInvPi;
  s.eta = 1.f;
  s.lobes = b.lobes;
  return b.reflectance * kInvPi * s.wo.z;
}

// ---------------------------------------------------------------- rough_metal (rough_metal.cpp:20-176)
__device__ inline f3 metal_eval(const BsdfRec& b, bool quirks, f3 wi, f3 wo, float& pdf) {
  pdf = 0.f;
  if (wi.z <= 0.f || wo.z <= 0.f) return splat(0.f);
  Microfacet mf = mf_of(b, quirks);
  f3 wh = normalize(wo + wi);
  f3 F = fresnel_conductor(dot(wi, wh), b.eta3, b.k3);
  float D = mf.D(wh), G = mf.G(wi, wo, wh);
  f3 brdf = absv(F * D * G) / (wi.z * wo.z * 4.f);
  pdf = mf.Pdf(wi, wh) / (4.f * dot(wo, wh));
  return mul(b.reflectance, brdf) * wo.z;
}
__device__ inline f3 metal_sample(const BsdfRec& b, bool quirks, f3 wi, float u1, float u2, BsdfSample& s) {
  s.pdf = 0.f;
  if (wi.z <= 0.f) return splat(0.f);
  Microfacet mf = mf_of(b, quirks);
  float pdf;
  f3 m = mf.Sample(wi, u1, u2, pdf);
  s.wo = reflect_wh(wi, m);
  s.eta = 1.f;
  s.lobes = b.lobes;
  if (pdf <= 0.f || s.wo.z <= 0.f) {
    s.pdf = 0.f;
    return splat(0.f);
  }
  s.pdf = pdf / (4.f * dot(s.wo, m));
  float D = mf.D(m), G = mf.G(wi, s.wo, m);
  f3 F = fresnel_conductor(dot(wi, m), b.eta3, b.k3);
  f3 brdf = (F * D * G) / (4.f * wi.z * s.wo.z);
  return mul(b.reflectance, brdf) * s.wo.z;
}

// ---------------------------------------------------------------- rough_glass (rough_glass.cpp:22-221)
__device__ inline f3 glass_eval(const BsdfRec& b, bool quirks, f3 wi, f3 wo, float& pdf) {
  pdf = 0.f;
  float ci = wi.z, co = wo.z;
  if (ci == 0.f) return splat(0.f);
  Microfacet mf = mf_of(b, quirks);
  bool refl = ci * co > 0.f;
  float e = ci > 0.f ? b.eta : b.inv_eta;
  float ie = ci > 0.f ? b.inv_eta : b.eta;
  f3 m = normalize(wi + wo * (refl ? 1.f : e));
  m = mulsign(m, m.z);
  float D = mf.D(m);
  float F = fresnel_dielectric(dot(wi, m), b.eta).r;
  float G = mf.G(wi, wo, m);
  float idm = dot(wi, m), odm = dot(wo, m);
  f3 val;
  if (refl) {
    val = b.reflectance * fabsf(F * D * G / (ci * 4.f));
  } else {
    float scale = sqr(ie);
    val = b.transmittance * fabsf((scale * (1.f - F) * D * G * e * e * idm * odm) / (ci * sqr(idm + e * odm)));
  }
  if (idm * ci <= 0.f || odm * co <= 0.f) {
    pdf = 0.f;
  } else {
    float dwhdwo = refl ? 1.f / (odm * 4.f) : ((e * e * odm) / sqr(idm + e * odm));
    float p = mf.Pdf(mulsign(wi, ci), m);
    p *= refl ? F : (1.f - F);
    pdf = fabsf(p * dwhdwo);
  }
  return val;
}
__device__ inline f3 glass_sample(const BsdfRec& b, bool quirks, f3 wi, float lobe, float u1, float u2,
                                  BsdfSample& s) {
  s.pdf = 0.f;
  float ci = wi.z;
  if (ci == 0.f) return splat(0.f);
  Microfacet mf = mf_of(b, quirks);
  float mpdf;
  f3 wh = mf.Sample(mulsign(wi, ci), u1, u2, mpdf);
  if (mpdf == 0.f) return splat(0.f);
  Dielectric fd = fresnel_dielectric(dot(wi, wh), b.eta);
  bool selR = lobe <= fd.r;
  s.wo = selR ? reflect_wh(wi, wh) : refract_wh(wi, wh, fd.cos_t, fd.eta_ti);
  bool isR = wi.z * s.wo.z > 0.f;
  s.eta = isR ? 1.f : fd.eta_it;
  s.lobes = LOBE_GLOSSY | (isR ? LOBE_REFLECTION : LOBE_TRANSMISSION);
  float pdf;
  f3 f = glass_eval(b, quirks, wi, s.wo, pdf);
  s.pdf = pdf;
  if (pdf <= 0.f) {
    s.pdf = 0.f;
    return splat(0.f);
  }
  return f;
}

// ---------------------------------------------------------------- disney (disney.cpp:28-613)
namespace disney {
__device__ inline float schlick_w(float c) {
  float m = fminf(fmaxf(1.f - c, 0.f), 1.f);
  return sqr(sqr(m)) * m;
}
__device__ inline float schlick_w_of(float cosThetaI, float eta) {
  bool out = cosThetaI >= 0.f;
  float rcpEta = 1.f / eta;
  float etaIT = out ? eta : rcpEta, etaTI = out ? rcpEta : eta;
  float sqrCosT = fmaf(-fmaf(-cosThetaI, cosThetaI, 1.f), sqr(etaTI), 1.f);
  float cosT = safe_sqrt(sqrCosT);
  return etaIT > 1.f ? schlick_w(fabsf(cosThetaI)) : schlick_w(cosT);
}
__device__ inline bool mac_mic(f3 m, f3 wi, f3 wo, float ci, bool refl) {
  if (refl) return dot(wi, mulsign(m, ci)) > 0.f && dot(wo, mulsign(m, ci)) > 0.f;
  return dot(wi, mulsign(m, ci)) > 0.f && dot(wo, ci >= 0.f ? -m : m) > 0.f;
}
__device__ inline float gtr1_eval(float alpha, f3 m) {
  float c2 = sqr(m.z), a2 = sqr(alpha);
  float r = (a2 - 1.f) / (kPi * logf(a2) * (1.f + (a2 - 1.f) * c2));
  return r * m.z > 1e-20f ? r : 0.f;
}
__device__ inline float gtr1_pdf(float alpha, f3 m) { return m.z < 0.f ? 0.f : m.z * gtr1_eval(alpha, m); }
__device__ inline f3 gtr1_sample(float alpha, float u1, float u2) {
  float a = 2.f * kPi * u1;
  float a2 = sqr(alpha);
  float c2 = (1.f - powf(a2, 1.f - u2)) / (1.f - a2);
  float st = sqrtf(fmaxf(0.f, 1.f - c2)), ct = sqrtf(fmaxf(0.f, c2));
  return mk(cosf(a) * st, sinf(a) * st, ct);
}
__device__ inline float smith_ggx1(f3 v, f3 wh, float alpha) {
  float a2 = sqr(alpha);
  float ct = fabsf(v.z), c2 = sqr(ct);
  float t2 = (1.f - c2) / c2;
  float r = 2.f / (1.f + sqrtf(1.f + a2 * t2));
  if (v.z == 1.f) r = 1.f;
  if (dot(v, wh) * v.z <= 0.f) r = 0.f;
  return r;
}
struct Params {
  float st, mt, cc, sh, fl, brdf, bsdf;
  bool hasST, hasCC, hasSheen, hasMetal, hasTint, hasSheenTint, hasAniso, hasFlat;
};
__device__ inline Params params(const BsdfRec& b) {
  Params p;
  uint32_t f = b.disney_flags;
  p.hasAniso = f & RAD_DISNEY_HAS_ANISOTROPIC;
  p.hasST = f & RAD_DISNEY_HAS_SPEC_TRANS;
  p.hasSheen = f & RAD_DISNEY_HAS_SHEEN;
  p.hasSheenTint = f & RAD_DISNEY_HAS_SHEEN_TINT;
  p.hasFlat = f & RAD_DISNEY_HAS_FLATNESS;
  p.hasTint = f & RAD_DISNEY_HAS_SPEC_TINT;
  p.hasMetal = f & RAD_DISNEY_HAS_METALLIC;
  p.hasCC = f & RAD_DISNEY_HAS_CLEARCOAT;
  p.st = p.hasST ? b.spec_trans : 0.f;
  p.mt = p.hasMetal ? b.metallic : 0.f;
  p.cc = p.hasCC ? b.clearcoat : 0.f;
  p.sh = p.hasSheen ? b.sheen : 0.f;
  p.fl = p.hasFlat ? b.flatness : 0.f;
  p.brdf = (1.f - p.mt) * (1.f - p.st);
  p.bsdf = (1.f - p.mt) * p.st;
  return p;
}
__device__ inline Microfacet spec_dist(const BsdfRec& b, const Params& p, bool quirks) {
  float r2 = sqr(b.roughness);
  float ax, ay;
  if (!p.hasAniso) {
    ax = ay = fmaxf(0.001f, r2);
  } else {
    float aspect = sqrtf(1.f - 0.9f * b.anisotropic);
    ax = fmaxf(0.001f, r2 / aspect);
    ay = fmaxf(0.001f, r2 * aspect);
  }
  return Microfacet{ax, ay, false, true, quirks};
}
__device__ inline void lobe_probs(const BsdfRec& b, const Params& p, bool front, float fr, float& ps, float& pt,
                                  float& pc, float& pd) {
  ps = front ? b.spec_rate * (1.f - p.bsdf * (1.f - fr)) : fr;
  pt = p.hasST ? (front ? b.spec_rate * p.bsdf * (1.f - fr) : (1.f - fr)) : 0.f;
  pc = p.hasCC ? (front ? 0.25f * p.cc * b.cc_rate : 0.f) : 0.f;
  pd = front ? p.brdf * b.diff_rate : 0.f;
}
}  // namespace disney

__device__ inline float disney_pdf(const BsdfRec& b, bool quirks, f3 wi, f3 wo) {
  using namespace disney;
  float ci = wi.z;
  if (ci == 0.f) return 0.f;
  Params p = params(b);
  bool front = ci > 0.f;
  float etaPath = front ? b.eta : 1.f / b.eta;
  float co = wo.z;
  bool refl = ci * co > 0.f, refr = ci * co < 0.f;
  f3 wh = normalize(wi + wo * (refl ? 1.f : etaPath));
  wh = mulsign(wh, wh.z);
  Microfacet sd = spec_dist(b, p, quirks);
  float fr = fresnel_dielectric(dot(wi, wh), b.eta).r;
  float ps, pt, pc, pd;
  lobe_probs(b, p, front, fr, ps, pt, pc, pd);
  float rt = 1.f / (ps + pt + pc + pd);
  ps *= rt;
  pt *= rt;
  pc *= rt;
  pd *= rt;
  float jac;
  if (p.hasST) {
    float idh = dot(wi, wh), odh = dot(wo, wh);
    jac = fabsf(refl ? 1.f / (4.f * odh) : (sqr(etaPath) * odh) / sqr(idh + etaPath * odh));
  } else {
    jac = fabsf(1.f / (4.f * dot(wo, wh)));
  }
  float pdf = 0.f;
  bool canR = mac_mic(wh, wi, wo, ci, true) && refl;
  if (canR) pdf += ps * sd.Pdf(mulsign(wi, ci), wh) * jac;
  if (refl) pdf += pd * (wo.z * kInvPi);
  if (p.hasST && mac_mic(wh, wi, wo, ci, false) && refr) pdf += pt * sd.Pdf(mulsign(wi, ci), wh) * jac;
  if (p.hasCC && canR) pdf += pc * gtr1_pdf(lerp_(0.1f, 0.001f, b.clearcoat_gloss), wh) * jac;
  return pdf;
}

__device__ inline f3 disney_f(const BsdfRec& b, bool quirks, f3 wi, f3 wo) {
  using namespace disney;
  float ci = wi.z;
  if (ci == 0.f) return splat(0.f);
  Params p = params(b);
  float co = wo.z;
  bool refl = ci * co > 0.f, refr = ci * co < 0.f;
  bool front = ci > 0.f;
  float invEta = 1.f / b.eta;
  float etaPath = front ? b.eta : invEta;
  float invEtaPath = front ? invEta : b.eta;
  Microfacet sd = spec_dist(b, p, quirks);
  f3 wh = normalize(wi + wo * (refl ? 1.f : etaPath));
  wh = mulsign(wh, wh.z);
  float fr = fresnel_dielectric(dot(wi, wh), b.eta).r;
  bool canR = mac_mic(wh, wi, wo, ci, true), canT = mac_mic(wh, wi, wo, ci, false);
  bool lobeSR = refl && canR && fr > 0.f;
  bool lobeCC = p.hasCC && p.cc > 0.f && refl && canR && front;
  bool lobeST = p.hasST && p.bsdf > 0.f && refr && canT && fr < 1.f;
  bool lobeD = p.brdf > 0.f && refl && front;
  bool lobeSh = p.hasSheen && p.sh > 0.f && refl && (1.f - p.mt) > 0.f && front;
  float D = sd.D(wh), G = sd.G(wi, wo, wh);
  f3 val = splat(0.f);
  f3 bc = b.base_color;
  if (lobeSR) {
    float lum = p.hasTint ? luminance(bc) : 1.f;
    float tint = p.hasTint ? b.spec_tint : 0.f;
    float cosI = dot(wi, wh);
    // DisneyFresnel (disney.cpp:568-593)
    float etaIT = cosI >= 0.f ? b.eta : 1.f / b.eta;
    f3 sf = splat(0.f);
    float w = schlick_w_of(cosI, b.eta);
    if (p.hasMetal) sf += mk(w + bc.x * (1.f - w), w + bc.y * (1.f - w), w + bc.z * (1.f - w)) * p.mt;
    if (p.hasTint) {
      f3 ct = lum > 0.f ? bc / lum : splat(1.f);
      f3 F0 = ct * sqr((etaIT - 1.f) / (etaIT + 1.f));
      sf += mk(w + F0.x * (1.f - w), w + F0.y * (1.f - w), w + F0.z * (1.f - w)) * ((1.f - p.mt) * tint);
    }
    f3 front_f = splat((1.f - p.mt) * (1.f - tint) * fr) + sf;
    f3 F = front ? front_f : splat(p.bsdf * fr);
    val += F * D * G / (4.f * fabsf(ci));
  }
  if (p.hasST && lobeST) {
    float scale = sqr(invEtaPath);
    float idh = dot(wi, wh), odh = dot(wo, wh);
    float w = p.bsdf * fabsf((scale * (1.f - fr) * D * G * etaPath * etaPath * idh * odh) / (ci * sqr(idh + etaPath * odh)));
    val += sqrtv(bc) * w;
  }
  if (p.hasCC && lobeCC) {
    float ccF = lerp_(schlick_w_of(dot(wi, wh), b.eta), 1.f, 0.04f);
    float ccD = gtr1_eval(lerp_(0.1f, 0.001f, b.clearcoat_gloss), wh);
    float ccG = smith_ggx1(wi, wh, 0.25f) * smith_ggx1(wo, wh, 0.25f);
    val += splat((p.cc * 0.25f) * ccF * ccD * ccG * fabsf(co));
  }
  if (lobeD) {
    float Fo = schlick_w(fabsf(co)), Fi = schlick_w(fabsf(ci));
    float fsDiff = (1.f - 0.5f * Fi) * (1.f - 0.5f * Fo);
    float cosD = dot(wh, wo);
    float Rr = 2.f * b.roughness * sqr(cosD);
    float fsRetro = Rr * (Fo + Fi + Fo * Fi * (Rr - 1.f));
    float k;
    if (p.hasFlat) {
      float Fss90 = Rr / 2.f;
      float Fss = lerp_(1.f, Fss90, Fo) * lerp_(1.f, Fss90, Fi);
      float fsSSS = 1.25f * (Fss * (1.f / (fabsf(co) + fabsf(ci)) - 0.5f) + 0.5f);
      k = lerp_(fsDiff + fsRetro, fsSSS, p.fl);
    } else {
      k = fsDiff + fsRetro;
    }
    val += bc * (p.brdf * fabsf(co)) * kInvPi * k;
    if (p.hasSheen && lobeSh) {
      float Fd = schlick_w(fabsf(cosD));
      if (p.hasSheenTint) {
        float lum = luminance(bc);
        f3 ct = lum > 0.f ? bc / lum : splat(1.f);
        f3 cs = splat(1.f) + (ct - splat(1.f)) * b.sheen_tint;
        val += cs * (p.sh * (1.f - p.mt) * Fd * fabsf(co));
      } else {
        val += splat(p.sh * (1.f - p.mt) * Fd * fabsf(co));
      }
    }
  }
  return val;
}

__device__ inline f3 disney_sample(const BsdfRec& b, bool quirks, f3 wi, float lobe, float u1, float u2,
                                   BsdfSample& s) {
  using namespace disney;
  s.pdf = 0.f;
  s.eta = 1.f;
  float ci = wi.z;
  if (ci == 0.f) return splat(0.f);
  Params p = params(b);
  float bsdfW = p.hasST ? p.bsdf : 0.f;
  bool front = ci > 0.f;
  Microfacet sd = spec_dist(b, p, quirks);
  float mpdf;
  f3 mS = sd.Sample(mulsign(wi, ci), u1, u2, mpdf);
  Dielectric fd = fresnel_dielectric(dot(wi, mS), b.eta);
  if (!front && bsdfW <= 0.f) return splat(0.f);
  float ps, pt, pc, pd;
  lobe_probs(b, p, front, fd.r, ps, pt, pc, pd);
  float rt = 1.f / (ps + pt + pc + pd);
  pt *= rt;
  pc *= rt;
  pd *= rt;
  float cur = 0.f;
  bool sD = lobe < pd;
  cur += pd;
  bool sC = p.hasCC && lobe >= cur && lobe < cur + pc;
  cur += pc;
  bool sT = p.hasST && lobe >= cur && lobe < cur + pt;
  cur += pt;
  bool sR = lobe >= cur;
  if (sR) {
    s.wo = reflect_wh(wi, mS);
    s.lobes = LOBE_REFLECTION | LOBE_GLOSSY;
    if (!(mac_mic(mS, wi, s.wo, ci, true) && ci * s.wo.z > 0.f)) return splat(0.f);
  }
  if (sT) {
    s.wo = refract_wh(wi, mS, fd.cos_t, fd.eta_ti);
    s.lobes = LOBE_TRANSMISSION | LOBE_GLOSSY;
    s.eta = fd.eta_it;
    if (!(mac_mic(mS, wi, s.wo, ci, false) && ci * s.wo.z < 0.f)) return splat(0.f);
  }
  if (p.hasCC && sC) {
    f3 mC = gtr1_sample(lerp_(0.1f, 0.001f, b.clearcoat_gloss), u1, u2);
    s.wo = reflect_wh(wi, mC);
    s.lobes = LOBE_REFLECTION | LOBE_GLOSSY;
    if (!(mac_mic(mC, wi, s.wo, ci, true) && ci * s.wo.z > 0.f)) return splat(0.f);
  }
  if (sD) {
    float r = sqrtf(u1), ang = 2.f * kPi * u2;
    float x = r * cosf(ang), y = r * sinf(ang);
    s.wo = mk(x, y, safe_sqrt(1.f - sqr(x) - sqr(y)));
    s.lobes = LOBE_REFLECTION | LOBE_DIFFUSE;
    if (!(ci * s.wo.z > 0.f)) return splat(0.f);
  }
  s.pdf = disney_pdf(b, quirks, wi, s.wo);
  if (s.pdf <= 0.f) {
    s.pdf = 0.f;
    return splat(0.f);
  }
  return disney_f(b, quirks, wi, s.wo);
}

// ---------------------------------------------------------------- plastic (plastic.cpp:18-132)
// The diffuse term under the dielectric coating, shared with rough_plastic:
// diffuse / pi (1 / eta^2) (1 - Fi) (1 - Fo) cos_o / (1 - Fdr), or the
// per-channel 1 / (1 - diffuse Fdr) with "no_linear".
__device__ inline f3 plastic_diffuse(const BsdfRec& b, float fi, float fo, float co) {
  f3 num = b.diffuse3 * kInvPi * b.inv_eta2 * (1.f - fi) * (1.f - fo) * co;
  if (b.nonlinear)
    return mk(num.x * (1.f / (1.f - b.diffuse3.x * b.fdr_int)), num.y * (1.f / (1.f - b.diffuse3.y * b.fdr_int)),
              num.z * (1.f / (1.f - b.diffuse3.z * b.fdr_int)));
  return num / (1.f - b.fdr_int);
}
__device__ inline float plastic_prob_spec(const BsdfRec& b, float fi) {
  float ps = fi * b.spec_weight, pd = (1.f - fi) * (1.f - b.spec_weight);
  return ps / (ps + pd);
}
__device__ inline f3 cosine_dir(float u1, float u2) {  // SquareToCosineHemisphere (warp.cpp:33-50)
  float r = sqrtf(u1), ang = 2.f * kPi * u2;
  float x = r * cosf(ang), y = r * sinf(ang);
  return mk(x, y, safe_sqrt(1.f - sqr(x) - sqr(y)));
}
__device__ inline f3 plastic_eval(const BsdfRec& b, f3 wi, f3 wo, float& pdf) {
  pdf = 0.f;
  if (wi.z <= 0.f || wo.z <= 0.f) return splat(0.f);
  float fi = fresnel_dielectric(wi.z, b.eta).r, fo = fresnel_dielectric(wo.z, b.eta).r;
  float ps = fi * b.spec_weight, pd = (1.f - fi) * (1.f - b.spec_weight);
  pdf = wo.z * kInvPi * (pd / (pd + ps));
  return plastic_diffuse(b, fi, fo, wo.z);
}
__device__ inline f3 plastic_sample(const BsdfRec& b, f3 wi, float lobe, float u1, float u2, BsdfSample& s) {
  s.pdf = 0.f;
  if (wi.z <= 0.f) return splat(0.f);
  float fi = fresnel_dielectric(wi.z, b.eta).r;
  float ps = plastic_prob_spec(b, fi);
  s.eta = 1.f;
  if (lobe < ps) {
    s.wo = mk(-wi.x, -wi.y, wi.z);
    s.pdf = ps;
    s.lobes = LOBE_REFLECTION | LOBE_DELTA;
    return b.spec3 * fi;
  }
  s.wo = cosine_dir(u1, u2);
  s.pdf = s.wo.z * kInvPi * (1.f - ps);
  s.lobes = LOBE_REFLECTION | LOBE_DIFFUSE;
  return plastic_diffuse(b, fi, fresnel_dielectric(s.wo.z, b.eta).r, s.wo.z);
}

// ---------------------------------------------------------------- rough_plastic (rough_plastic.cpp:18-250)
__device__ inline Microfacet mf_iso(const BsdfRec& b, bool quirks) {
  return Microfacet{b.alpha_u, b.alpha_u, b.beckmann != 0, b.sample_visible != 0, quirks};
}
__device__ inline f3 rough_plastic_eval(const BsdfRec& b, bool quirks, f3 wi, f3 wo, float& pdf) {
  pdf = 0.f;
  if (wi.z <= 0.f || wo.z <= 0.f) return splat(0.f);
  Microfacet mf = mf_iso(b, quirks);
  f3 wh = normalize(wo + wi);
  float F = fresnel_dielectric(dot(wi, wh), b.eta).r;
  f3 spec = b.spec3 * mf.D(wh) * mf.G(wi, wo, wh) * F / (4.f * wi.z);
  float fo = fresnel_dielectric(dot(wo, wh), b.eta).r;
  float ps = plastic_prob_spec(b, fresnel_dielectric(wi.z, b.eta).r);
  pdf = mf.Pdf(wi, wh) / (4.f * dot(wo, wh)) * ps + wo.z * kInvPi * (1.f - ps);
  return spec + plastic_diffuse(b, F, fo, wo.z);
}
__device__ inline f3 rough_plastic_sample(const BsdfRec& b, bool quirks, f3 wi, float lobe, float u1, float u2,
                                          BsdfSample& s) {
  s.pdf = 0.f;
  if (wi.z <= 0.f) return splat(0.f);
  float ps = plastic_prob_spec(b, fresnel_dielectric(wi.z, b.eta).r);
  s.eta = 1.f;
  if (lobe < ps) {
    Microfacet mf = mf_iso(b, quirks);
    float mpdf;
    f3 m = mf.Sample(wi, u1, u2, mpdf);
    if (mpdf <= 0.f) return splat(0.f);
    s.wo = reflect_wh(wi, m);
    s.lobes = LOBE_GLOSSY | LOBE_REFLECTION;
  } else {
    s.wo = cosine_dir(u1, u2);
    s.lobes = LOBE_DIFFUSE | LOBE_REFLECTION;
  }
  float pdf;
  f3 f = rough_plastic_eval(b, quirks, wi, s.wo, pdf);
  if (!(pdf > 0.f)) return splat(0.f);
  s.pdf = pdf;
  return f;
}

// ---------------------------------------------------------------- perfect_mirror / perfect_glass
// (perfect_mirror.cpp:24-43, perfect_glass.cpp:26-65): delta lobes only, so
// Eval and Pdf are zero.
__device__ inline f3 mirror_sample(const BsdfRec& b, f3 wi, BsdfSample& s) {
  s.pdf = 0.f;
  if (wi.z <= 0.f) return splat(0.f);
  s.wo = mk(-wi.x, -wi.y, wi.z);
  s.eta = 1.f;
  s.pdf = 1.f;
  s.lobes = LOBE_DELTA | LOBE_REFLECTION;
  return mul(b.reflectance, fresnel_conductor(wi.z, b.eta3, b.k3));
}
__device__ inline f3 perfect_glass_sample(const BsdfRec& b, f3 wi, float lobe, BsdfSample& s) {
  Dielectric fd = fresnel_dielectric(wi.z, b.eta);
  float ri = fd.r, ti = 1.f - ri;
  if (lobe <= ri) {
    s.wo = mk(-wi.x, -wi.y, wi.z);
    s.eta = 1.f;
    s.pdf = ri;
    s.lobes = LOBE_DELTA | LOBE_REFLECTION;
    return b.reflectance * ri;
  }
  s.wo = mk(-fd.eta_ti * wi.x, -fd.eta_ti * wi.y, fd.cos_t);
  s.eta = fd.eta_it;
  s.pdf = ti;
  s.lobes = LOBE_DELTA | LOBE_TRANSMISSION;
  return b.transmittance * ti * sqr(fd.eta_ti);
}

// ---------------------------------------------------------------- dispatch over the base BSDFs
__device__ inline f3 bsdf_eval(const BsdfRec& b, bool quirks, f3 wi, f3 wo, float& pdf) {
  switch (b.type) {
    case BSDF_DIFFUSE:
      return diffuse_eval(b, wi, wo, pdf);
    case BSDF_ROUGH_METAL:
      return metal_eval(b, quirks, wi, wo, pdf);
    case BSDF_ROUGH_GLASS:
      return glass_eval(b, quirks, wi, wo, pdf);
    case BSDF_DISNEY:
      pdf = disney_pdf(b, quirks, wi, wo);
      return disney_f(b, quirks, wi, wo);
    case BSDF_PLASTIC:
      return plastic_eval(b, wi, wo, pdf);
    case BSDF_ROUGH_PLASTIC:
      return rough_plastic_eval(b, quirks, wi, wo, pdf);
    default:  // perfect mirror / glass
      pdf = 0.f;
      return splat(0.f);
  }
}

__device__ inline f3 bsdf_sample(const BsdfRec& b, bool quirks, f3 wi, float lobe, float u1, float u2, BsdfSample& s) {
  switch (b.type) {
    case BSDF_DIFFUSE:
      return diffuse_sample(b, wi, u1, u2, s);
    case BSDF_ROUGH_METAL:
      return metal_sample(b, quirks, wi, u1, u2, s);
    case BSDF_ROUGH_GLASS:
      return glass_sample(b, quirks, wi, lobe, u1, u2, s);
    case BSDF_DISNEY:
      return disney_sample(b, quirks, wi, lobe, u1, u2, s);
    case BSDF_PLASTIC:
      return plastic_sample(b, wi, lobe, u1, u2, s);
    case BSDF_ROUGH_PLASTIC:
      return rough_plastic_sample(b, quirks, wi, lobe, u1, u2, s);
    case BSDF_PERFECT_MIRROR:
      return mirror_sample(b, wi, s);
    default:
      return perfect_glass_sample(b, wi, lobe, s);
  }
}

// ---------------------------------------------------------------- two_side / mask wrappers
// (two_side.cpp:60-137, mask.cpp:17-53).  At most two wrappers enclose a
// base BSDF (the scene loader checks).  two_side evaluates the lower
// hemisphere through the mirrored directions; mask scales by its opacity and
// otherwise passes the path straight through (weight 1, pdf 1 - opacity, no
// lobe type, as the reference records it).
__device__ inline f3 bsdf_eval_w(const BsdfRec* all, const BsdfRec& top, bool quirks, f3 wi, f3 wo, float& pdf) {
  const BsdfRec* b = &top;
  float scale = 1.f;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (b->type == BSDF_MASK) {
      scale *= b->opacity;
      b = &all[b->nested];
    } else if (b->type == BSDF_TWO_SIDE) {
      if (wi.z == 0.f) {
        pdf = 0.f;
        return splat(0.f);
      }
      if (wi.z < 0.f) {
        wi.z = -wi.z;
        wo.z = -wo.z;
      }
      b = &all[b->nested];
    }
  }
  f3 f = bsdf_eval(*b, quirks, wi, wo, pdf);
  pdf *= scale;
  return f * scale;
}

__device__ inline f3 bsdf_sample_w(const BsdfRec* all, const BsdfRec& top, bool quirks, f3 wi, float lobe, float u1,
                                   float u2, BsdfSample& s) {
  const BsdfRec* b = &top;
  float scale = 1.f;
  bool flip = false;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (b->type == BSDF_MASK) {
      const float op = b->opacity;
      if (!(lobe < op)) {
        s.wo = -wi;
        if (flip) s.wo.z = -s.wo.z;
        s.pdf = (1.f - op) * scale;
        s.eta = 1.f;
        s.lobes = 0u;
        return splat(scale);
      }
      lobe /= op;
      scale *= op;
      b = &all[b->nested];
    } else if (b->type == BSDF_TWO_SIDE) {
      if (wi.z == 0.f) {
        s.pdf = 0.f;
        return splat(0.f);
      }
      if (wi.z < 0.f) {
        wi.z = -wi.z;
        flip = !flip;
      }
      b = &all[b->nested];
    }
  }
  f3 f = bsdf_sample(*b, quirks, wi, lobe, u1, u2, s);
  if (flip) s.wo.z = -s.wo.z;
  s.pdf *= scale;
  return f * scale;
}

// Material-specialised dispatch for kernels that shade one BSDF type
// (MAT = BSDF_DIFFUSE .. BSDF_DISNEY); any other MAT handles every type,
// wrappers included, at run time.
template <int MAT>
__device__ __forceinline__ f3 bsdf_eval_t(const BsdfRec* all, const BsdfRec& b, bool quirks, f3 wi, f3 wo, float& pdf) {
  if constexpr (MAT == BSDF_DIFFUSE) {
    return diffuse_eval(b, wi, wo, pdf);
  } else if constexpr (MAT == BSDF_ROUGH_METAL) {
    return metal_eval(b, quirks, wi, wo, pdf);
  } else if constexpr (MAT == BSDF_ROUGH_GLASS) {
    return glass_eval(b, quirks, wi, wo, pdf);
  } else if constexpr (MAT == BSDF_DISNEY) {
    pdf = disney_pdf(b, quirks, wi, wo);
    return disney_f(b, quirks, wi, wo);
  } else {
    return bsdf_eval_w(all, b, quirks, wi, wo, pdf);
  }
}

template <int MAT>
__device__ __forceinline__ f3 bsdf_sample_t(const BsdfRec* all, const BsdfRec& b, bool quirks, f3 wi, float lobe,
                                            float u1, float u2, BsdfSample& s) {
  if constexpr (MAT == BSDF_DIFFUSE) {
    return diffuse_sample(b, wi, u1, u2, s);
  } else if constexpr (MAT == BSDF_ROUGH_METAL) {
    return metal_sample(b, quirks, wi, u1, u2, s);
  } else if constexpr (MAT == BSDF_ROUGH_GLASS) {
    return glass_sample(b, quirks, wi, lobe, u1, u2, s);
  } else if constexpr (MAT == BSDF_DISNEY) {
    return disney_sample(b, quirks, wi, lobe, u1, u2, s);
  } else {
    return bsdf_sample_w(all, b, quirks, wi, lobe, u1, u2, s);
  }
}

}  // namespace rad
