// Device side of path guiding: the one-blob input encoding, the decoded
// von Mises-Fisher mixture (pdf and sampling), and the rule for which
// vertices are guided.
#pragma once

#include "guide.h"
#include "shading_dev.h"

namespace rad {
namespace guide {

// Scalars in [0, 1] that the one-blob encoding spreads over bins:
// position (3), normal (polar, azimuth), outgoing direction (polar, azimuth).
struct FeatScalars {
  float s[7];
};

RAD_HD float dir_polar(f3 d) { return fminf(fmaxf(0.5f * (d.y + 1.f), 0.f), 1.f); }
__device__ inline float dir_azimuth(f3 d) { return fminf(fmaxf(atan2f(d.z, d.x) * (0.5f / kPi) + 0.5f, 0.f), 1.f); }

__device__ inline FeatScalars feat_scalars(const float* f9, const float lo[3], const float inv_ext[3]) {
  FeatScalars r;
  for (int i = 0; i < 3; i++) r.s[i] = fminf(fmaxf((f9[i] - lo[i]) * inv_ext[i], 0.f), 1.f);
  f3 n = mk(f9[3], f9[4], f9[5]), wi = mk(f9[6], f9[7], f9[8]);
  r.s[3] = dir_polar(n);
  r.s[4] = dir_azimuth(n);
  r.s[5] = dir_polar(wi);
  r.s[6] = dir_azimuth(wi);
  return r;
}

// One-blob value of input feature 16 c + r (0 <= r < 16): a unit-width
// triangle kernel centred on the scalar's position among the bins.
// Features come in 16-wide chunks: chunks 0-2 are the 16 bins of position
// x, y, z; chunks 3 and 4 the 8 bins of (normal polar, azimuth) and
// (direction polar, azimuth).  Callers pass the chunk as a constant, so the
// scalar is picked without an indexed (scratch) access.
static_assert(kPosBins == 16 && kDirBins == 8 && kIn == 80, "one-blob chunk layout");
__device__ __forceinline__ float one_blob(const FeatScalars& fs, int c, int r) {
  float sv;
  int bins, bin;
  if (c < 3) {
    sv = c == 0 ? fs.s[0] : c == 1 ? fs.s[1] : fs.s[2];
    bins = kPosBins;
    bin = r;
  } else {
    const bool hi = r >= kDirBins;
    sv = c == 3 ? (hi ? fs.s[4] : fs.s[3]) : (hi ? fs.s[6] : fs.s[5]);
    bins = kDirBins;
    bin = r & (kDirBins - 1);
  }
  float x = sv * float(bins) - (float(bin) + 0.5f);
  return fmaxf(0.f, 1.f - fabsf(x));
}

// Hardware transcendentals (v_exp_f32, v_log_f32, v_rcp_f32, v_rsq_f32,
// v_sin_f32, v_cos_f32; about 1 ulp) for the mixture's own arithmetic: its
// density and its sampler agree to a relative ~1e-6, far below the noise of
// the estimate, and shading evaluates the mixture several times per vertex.
__device__ __forceinline__ float hw_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }
__device__ __forceinline__ float hw_log(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }

// One lane's decoded mixture, held in LDS rather than registers (48 floats
// per lane would cost the shading kernels a wave per SIMD): lobe l of the
// lane is km[l * stride] (mean direction, kappa) and pw[l * stride]
// (weight, weighted vMF normalisation pi kappa / (2 pi (1 - e^{-2 kappa}))).
// A block of B lanes gives lane t the pointers km0 + t, pw0 + t and stride B,
// so a lobe read by a wave is one contiguous LDS row.
struct Mixture {
  float4* km;
  float2* pw;
  uint32_t stride;
};

__device__ __forceinline__ Mixture lane_mixture(float4* km0, float2* pw0, uint32_t lane, uint32_t lanes) {
  return Mixture{km0 + lane, pw0 + lane, lanes};
}

// Vector from vertex p to a lobe's target c + h * (its three mean outputs).
template <class T>
__device__ __forceinline__ f3 lobe_vector(const T* lobe, f3 p, f3 cen, f3 half) {
  return mk(fmaf(half.x, float(lobe[1]), cen.x) - p.x, fmaf(half.y, float(lobe[2]), cen.y) - p.y,
            fmaf(half.z, float(lobe[3]), cen.z) - p.z);
}

// raw: 40 network outputs (logit, target xyz, log kappa per lobe) of the
// vertex at p; cen, half: the lobe-target frame (GuideNet::center, half).
template <class T>
__device__ __forceinline__ void decode(const Mixture& m, const T* raw, f3 p, f3 cen, f3 half) {
  float e[kLobes];
  float mx = -kInf;
#pragma unroll
  for (int l = 0; l < kLobes; l++) mx = fmaxf(mx, float(raw[kRawPerLobe * l]));
  float sum = 0.f;
#pragma unroll
  for (int l = 0; l < kLobes; l++) {
    e[l] = hw_exp(float(raw[kRawPerLobe * l]) - mx);
    sum += e[l];
  }
  const float inv_sum = __builtin_amdgcn_rcpf(sum);
#pragma unroll
  for (int l = 0; l < kLobes; l++) {
    const float pi = e[l] * inv_sum;
    f3 v = lobe_vector(raw + kRawPerLobe * l, p, cen, half);
    const float len2 = sqnorm(v);
    f3 mu = len2 > 1e-12f ? v * __builtin_amdgcn_rsqf(len2) : mk(0.f, 1.f, 0.f);
    float kappa = hw_exp(fminf(fmaxf(float(raw[kRawPerLobe * l + 4]), -3.f), 6.f));
    m.km[l * m.stride] = make_float4(mu.x, mu.y, mu.z, kappa);
    // kappa >= e^-3, so 1 - e^{-2 kappa} >= 0.095 needs no expm1
    m.pw[l * m.stride] = make_float2(pi, pi * kappa * __builtin_amdgcn_rcpf((2.f * kPi) * (1.f - hw_exp(-2.f * kappa))));
  }
}

// vMF density on the sphere: kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (mu.w - 1)}.
__device__ __forceinline__ float vmf_pdf(f3 mu, float kappa, f3 w) {
  float norm = kappa / ((2.f * kPi) * (-expm1f(-2.f * kappa)));
  return norm * expf(kappa * (dot(mu, w) - 1.f));
}

// The normalisations are taken from the decode, so an evaluation costs one
// exponential per lobe.
__device__ __forceinline__ float mixture_pdf(const Mixture& m, f3 w) {
  float p = 0.f;
#pragma unroll
  for (int l = 0; l < kLobes; l++) {
    const float4 a = m.km[l * m.stride];
    p += m.pw[l * m.stride].y * hw_exp(a.w * (dot(mk(a.x, a.y, a.z), w) - 1.f));
  }
  return p;
}

// Pick a lobe with u0 by its weights, then sample it by inverting the CDF of
// the cosine to the lobe axis.
__device__ __forceinline__ f3 mixture_sample(const Mixture& m, float u0, float u1, float u2) {
  uint32_t pick = kLobes - 1;
  float c = 0.f;
#pragma unroll
  for (int l = 0; l < kLobes - 1; l++) {
    c += m.pw[l * m.stride].x;
    if (pick == kLobes - 1 && u0 < c) pick = l;
  }
  const float4 a = m.km[pick * m.stride];
  const f3 mu = mk(a.x, a.y, a.z);
  const float k = a.w;
  float w = 1.f + hw_log(fmaxf(u1 + (1.f - u1) * hw_exp(-2.f * k), 1e-30f)) * __builtin_amdgcn_rcpf(k);
  w = fminf(fmaxf(w, -1.f), 1.f);
  float r = safe_sqrt(1.f - w * w);
  Frame fr = frame_from_n(mu);
  // azimuth 2 pi u2: the hardware sine and cosine take revolutions
  return normalize(fr.to_world(mk(r * __builtin_amdgcn_cosf(u2), r * __builtin_amdgcn_sinf(u2), w)));
}

// A vertex is guided when its BSDF is reflective only (a guided refraction
// would need the BSDF's relative IOR), has no delta lobe, and its sampler's
// density equals its pdf: with the reference's visible-normal azimuth quirk
// on, only BSDFs that do not sample visible normals qualify.
// Glossy BSDFs below min_rough are left to BSDF sampling, which is already
// close to their product distribution.
__device__ __forceinline__ bool guidable(const BsdfRec& b, bool quirks, float min_rough) {
  if (b.lobes & (LOBE_DELTA | LOBE_TRANSMISSION)) return false;
  if (b.type > BSDF_DISNEY) return false;  // plastics carry a delta lobe; wrappers pass paths through
  if (b.type == BSDF_DIFFUSE) return true;
  const float rough = b.type == BSDF_DISNEY ? b.roughness * b.roughness : fminf(b.alpha_u, b.alpha_v);
  if (rough < min_rough) return false;
  if (!quirks) return true;
  return (b.type == BSDF_ROUGH_METAL || b.type == BSDF_ROUGH_GLASS) && !b.sample_visible;
}

// Geometric normal of a hit, for the network inputs.
__device__ inline f3 geometric_normal(const DevScene& s, f3 p, uint32_t prim, uint32_t geom) {
  const ShapeRec& sr = s.shapes[geom];
  if (sr.type == SHAPE_MESH) {
    const float4* T = s.tris + 6 * size_t(sr.tri_offset + prim);  // DevScene::tris
    const float4 t0 = T[0], t1 = T[1], t2 = T[2];
    f3 p0 = mk(t0.x, t0.y, t0.z), p1 = mk(t0.w, t1.x, t1.y), p2 = mk(t1.z, t1.w, t2.x);
    return normalize(cross(p1 - p0, p2 - p0));
  }
  if (sr.type == SHAPE_SPHERE) return normalize(p - sr.center);
  return sr.frame_n;
}

}  // namespace guide
}  // namespace rad
