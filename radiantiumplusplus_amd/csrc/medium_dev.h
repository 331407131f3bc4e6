// Participating media on the device: the heterogeneous grid lookup
// (volume/grid.cpp:41-103), delta-tracking distance sampling
// (medium.cpp:19-43) and the phase functions (henyey_greenstein.cpp:18-40,
// isotropic.cpp:16-24).
#pragma once

#include "shading_dev.h"

namespace rad {

__device__ inline float wrap_coord(float v, int wrap) {
  if (wrap == RAD_WRAP_REPEAT) return fminf(fmaxf(v - floorf(v), 0.f), 1.f);
  return fminf(fmaxf(v, 0.f), 1.f);
}

__device__ inline f3 grid_read(const float* g, int sx, int sy, int ch, int u, int v, int w) {
  size_t start = ((size_t(w) * size_t(sy) + size_t(v)) * size_t(sx) + size_t(u)) * size_t(ch);
  if (ch == 1) return splat(g[start]);
  return mk(g[start], g[start + 1], g[start + 2]);
}

// The reference's "trilinear" lookup: scaled by (size - 1); the neighbour of
// each axis is chosen by the half-voxel offset and weighted by |frac - 0.5|.
__device__ inline f3 grid_eval(const float* g, int sx, int sy, int sz, int ch, int wrap, f3 p) {
  float wx = wrap_coord(p.x, wrap), wy = wrap_coord(p.y, wrap), wz = wrap_coord(p.z, wrap);
  float fx = wx * float(sx - 1), fy = wy * float(sy - 1), fz = wz * float(sz - 1);
  int ix = int(fx), iy = int(fy), iz = int(fz);
  int ax = min(max(ix + ((fx > (float(ix) + 0.5f)) ? 1 : -1), 0), sx - 1);
  int ay = min(max(iy + ((fy > (float(iy) + 0.5f)) ? 1 : -1), 0), sy - 1);
  int az = min(max(iz + ((fz > (float(iz) + 0.5f)) ? 1 : -1), 0), sz - 1);
  float du = fminf(fabsf(fx - float(ix) - 0.5f), 1.f);
  float dv = fminf(fabsf(fy - float(iy) - 0.5f), 1.f);
  float dw = fminf(fabsf(fz - float(iz) - 0.5f), 1.f);
  f3 c000 = grid_read(g, sx, sy, ch, ix, iy, iz), c100 = grid_read(g, sx, sy, ch, ax, iy, iz);
  f3 c010 = grid_read(g, sx, sy, ch, ix, ay, iz), c001 = grid_read(g, sx, sy, ch, ix, iy, az);
  f3 c110 = grid_read(g, sx, sy, ch, ax, ay, iz), c011 = grid_read(g, sx, sy, ch, ix, ay, az);
  f3 c101 = grid_read(g, sx, sy, ch, ax, iy, az), c111 = grid_read(g, sx, sy, ch, ax, ay, az);
  f3 x00 = fmadd(c000, splat(1.f - du), c100 * du);
  f3 x10 = fmadd(c010, splat(1.f - du), c110 * du);
  f3 y0 = fmadd(x00, splat(1.f - dv), x10 * dv);
  f3 x01 = fmadd(c001, splat(1.f - du), c101 * du);
  f3 x11 = fmadd(c011, splat(1.f - du), c111 * du);
  f3 y1 = fmadd(x01, splat(1.f - dv), x11 * dv);
  return fmadd(y0, splat(1.f - dw), y1 * dw);
}

// ---------------------------------------------------------------- media
// A sampled medium interaction (interaction.h:102-111): T = kInf when the
// sampled distance lies beyond the ray segment inside the medium.
struct MediumSample {
  float T, minT;
  f3 p;
  f3 sig_s, sig_n, sig_t, comb;
};

__device__ inline f3 medium_sigma_t(const DevScene& s, const MediumRec& m, f3 p) {
  if (m.type == RAD_MEDIUM_HOMOGENEOUS) return m.sigma_t;
  if (!m.has_grid) return m.sigma_t * m.scale;  // heterogeneous.cpp:41-44
  f3 q = m.to_local.point(p);
  return grid_eval(s.grid_data + m.grid_offset, m.sx, m.sy, m.sz, m.channels, m.wrap, q) * m.scale;
}

// Medium::SampleInteraction (medium.cpp:19-43) with the homogeneous /
// heterogeneous IntersectAABB, GetMajorant and GetScatteringCoefficients.
__device__ inline MediumSample medium_sample(const DevScene& s, const MediumRec& m, f3 o, f3 d, float rayMaxT, float u,
                                             uint32_t ch) {
  MediumSample mi;
  bool active;
  float mint, maxt;
  if (m.type == RAD_MEDIUM_HOMOGENEOUS) {  // homogeneous.cpp:23-25
    active = true;
    mint = 0.f;
    maxt = kInf;
  } else {  // heterogeneous.cpp:66-79
    active = !is_zero(d) || (o.x >= m.bbox_lo.x && o.y >= m.bbox_lo.y && o.z >= m.bbox_lo.z && o.x <= m.bbox_hi.x &&
                             o.y <= m.bbox_hi.y && o.z <= m.bbox_hi.z);
    f3 r = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    f3 t1 = mul(m.bbox_lo - o, r), t2 = mul(m.bbox_hi - o, r);
    f3 a = mk(fminf(t1.x, t2.x), fminf(t1.y, t2.y), fminf(t1.z, t2.z));
    f3 b = mk(fmaxf(t1.x, t2.x), fmaxf(t1.y, t2.y), fmaxf(t1.z, t2.z));
    mint = fmaxf(fmaxf(a.x, a.y), a.z);
    maxt = fminf(fminf(b.x, b.y), b.z);
    active = active && maxt >= mint;
  }
  active = active && (isfinite(mint) || isfinite(maxt));
  if (!active) {
    mint = 0.f;
    maxt = kInf;
  }
  mint = fmaxf(0.f, mint);
  maxt = fminf(rayMaxT, maxt);
  mi.comb = m.type == RAD_MEDIUM_HOMOGENEOUS ? m.sigma_t : splat(m.max_density);
  float mc = ch == 0 ? mi.comb.x : (ch == 1 ? mi.comb.y : mi.comb.z);
  float t = mint + (-logf(1.f - u) / mc);
  mi.T = (active && t <= maxt) ? t : kInf;
  mi.p = d * t + o;
  mi.minT = mint;
  mi.sig_t = medium_sigma_t(s, m, mi.p);
  mi.sig_s = mul(mi.sig_t, m.albedo);
  mi.sig_n = m.type == RAD_MEDIUM_HOMOGENEOUS ? splat(0.f) : splat(m.max_density) - mi.sig_t;
  return mi;
}

RAD_HD float chan(f3 v, uint32_t c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }

// Medium::EvalTrAndPdf (medium.cpp:45-59); siT = kInf on a miss.
__device__ inline void medium_tr_pdf(const MediumSample& mi, float siT, f3& tr, f3& pdf) {
  float t = fminf(mi.T, siT) - mi.minT;
  if (t <= 0.f) {
    tr = pdf = splat(1.f);
    return;
  }
  tr = mk(expf(mi.comb.x * -t), expf(mi.comb.y * -t), expf(mi.comb.z * -t));
  pdf = siT < mi.T ? tr : mul(tr, mi.comb);
  if (tr.x == pdf.x && tr.y == pdf.y && tr.z == pdf.z) tr = pdf = splat(1.f);
}

RAD_HD float eval_hg(float g, float cosTheta) {  // henyey_greenstein.cpp:11-16
  float temp = 1.f + g * g + 2.f * g * cosTheta;
  return (1.f / (4.f * kPi)) * (1.f - g * g) / (temp * sqrtf(temp));
}

// Phase sampling in the frame of wi = -D (henyey_greenstein.cpp:18-34,
// isotropic.cpp:16-21).
__device__ inline f3 phase_sample(const MediumRec& m, f3 wi, float u1, float u2, float& pdf) {
  if (m.phase != RAD_PHASE_HG) {
    float z = fmaf(-2.f, u2, 1.f), r = safe_sqrt(fmaf(-z, z, 1.f)), a = 2.f * kPi * u1;
    pdf = 1.f / (4.f * kPi);
    return mk(r * cosf(a), r * sinf(a), z);
  }
  float g = m.g;
  float sqrTerm = (1.f - g * g) / (1.f - g + 2.f * g * u1);
  float cosTheta = (1.f + g * g - sqrTerm * sqrTerm) / (2.f * g);
  if (fabsf(g) < kFloat32Eps) cosTheta = 1.f - 2.f * u1;
  float sinTheta = safe_sqrt(1.f - cosTheta * cosTheta);
  float a = 2.f * kPi * u2;
  Frame fr = frame_from_n(wi);
  f3 wo = fr.to_world(mk(sinTheta * cosf(a), sinTheta * sinf(a), -cosTheta));
  pdf = eval_hg(g, -cosTheta);
  return wo;
}

__device__ inline float phase_eval(const MediumRec& m, f3 wi, f3 wo) {
  if (m.phase != RAD_PHASE_HG) return 1.f / (4.f * kPi);
  return eval_hg(m.g, dot(wo, wi));
}

}  // namespace rad
