// Device texture lookups: `bitmap` (nearest, the reference's bilinear,
// trilinear over the mip chain, EWA) and `chessboard`
// (texture/bitmap.cpp:24-247, texture/chessboard.cpp:33-55).
//
// Arithmetic is kept uncontracted so that texel selection (the integer
// truncations, the EWA window and its r^2 < 1 test) follows the same float
// values as the serial definition.
#pragma once

#include "scene_dev.h"

namespace rad {

#pragma float_control(push)
#pragma clang fp contract(off)

struct TexCoord {
  float u, v;
  float dudx, dvdx, dudy, dvdy;
};

__device__ inline f3 texel(const DevScene& s, const TexRec& t, uint32_t level, uint32_t x, uint32_t y) {
  const float* p = s.texels + t.off[level] + (size_t(y) * t.w[level] + x) * uint32_t(t.channels);
  return t.channels == 3 ? mk(p[0], p[1], p[2]) : mk(p[0], 0.f, 0.f);
}

// MipMap::Wrap for coordinates (bitmap.cpp:136-144)
__device__ inline float tex_wrap(float v, int32_t wrap) {
  if (wrap == RAD_WRAP_REPEAT) return fminf(fmaxf(v - floorf(v), 0.f), 1.f);
  return fminf(fmaxf(v, 0.f), 1.f);
}

// bitmap.cpp:192-206: the neighbour is on the side of the texel centre the
// coordinate falls; its index is formed in unsigned arithmetic and clamped
// (below texel 0 that is the row's last texel).
__device__ inline f3 tex_bilinear4(const DevScene& s, const TexRec& t, uint32_t l, uint32_t pu, uint32_t pv, float fu,
                                   float fv) {
  const uint32_t W = t.w[l], H = t.h[l];
  const uint32_t apu = min(pu + (fu > float(pu) + 0.5f ? 1u : 0xffffffffu), W - 1);
  const uint32_t apv = min(pv + (fv > float(pv) + 0.5f ? 1u : 0xffffffffu), H - 1);
  const float du = fminf(fabsf(fu - float(pu) - 0.5f), 1.f);
  const float dv = fminf(fabsf(fv - float(pv) - 0.5f), 1.f);
  const f3 a = texel(s, t, l, pu, pv), b = texel(s, t, l, apu, pv);
  const f3 c = texel(s, t, l, pu, apv), d = texel(s, t, l, apu, apv);
  return (a * (1.f - du) + b * du) * (1.f - dv) + (c * (1.f - du) + d * du) * dv;
}

__device__ inline f3 tex_nearest(const DevScene& s, const TexRec& t, float u, float v) {  // bitmap.cpp:168-178
  const float fu = tex_wrap(u, t.wrap) * float(t.w[0] - 1);
  const float fv = tex_wrap(v, t.wrap) * float(t.h[0] - 1);
  return texel(s, t, 0, uint32_t(fu), uint32_t(fv));
}

__device__ inline f3 tex_bilinear(const DevScene& s, const TexRec& t, uint32_t l, float u, float v) {  // :180-190
  const float fu = tex_wrap(u, t.wrap) * float(t.w[l] - 1);
  const float fv = tex_wrap(v, t.wrap) * float(t.h[l] - 1);
  return tex_bilinear4(s, t, l, uint32_t(fu), uint32_t(fv), fu, fv);
}

// bitmap.cpp:214-246: weights on the level's texel grid, values read from
// level 0 by nearest neighbour, as the reference does.
__device__ inline f3 tex_ewa(const DevScene& s, const TexRec& t, uint32_t l, float u, float v, float d0x, float d0y,
                             float d1x, float d1y) {
  const float Wl = float(t.w[l]), Hl = float(t.h[l]);
  const float stx = u * Wl, sty = v * Hl;
  const float e0x = d0x * Wl, e0y = d0y * Hl, e1x = d1x * Wl, e1y = d1y * Hl;
  float A = e0y * e0y + e1y * e1y + 1.f;
  float B = -2.f * (e0x * e0y + e1x * e1y);
  float C = e0x * e0x + e1x * e1x + 1.f;
  const float invF = 1.f / (A * C - B * B * 0.25f);
  A *= invF;
  B *= invF;
  C *= invF;
  const float det = -B * B + 4.f * A * C;
  const float invDet = 1.f / det;
  const float uSqrt = sqrtf(det * C), vSqrt = sqrtf(det * A);
  const int32_t s0 = int32_t(ceilf(stx - 2.f * invDet * uSqrt)), s1 = int32_t(floorf(stx + 2.f * invDet * uSqrt));
  const int32_t t0 = int32_t(ceilf(sty - 2.f * invDet * vSqrt)), t1 = int32_t(floorf(sty + 2.f * invDet * vSqrt));
  const float irx = 1.f / Wl, iry = 1.f / Hl;
  const float e2 = expf(-2.f);
  f3 sum = splat(0.f);
  float sumW = 0.f;
  for (int32_t it = t0; it <= t1; ++it) {
    const float tt = float(it) - sty;
    for (int32_t is = s0; is <= s1; ++is) {
      const float ss = float(is) - stx;
      const float r2 = A * ss * ss + B * ss * tt + C * tt * tt;
      if (r2 < 1.f) {
        const float wgt = expf(-2.f * r2) - e2;
        sum = sum + tex_nearest(s, t, float(is) * irx, float(it) * iry) * wgt;
        sumW += wgt;
      }
    }
  }
  return sum / sumW;
}

__device__ inline f3 tex_trilinear(const DevScene& s, const TexRec& t, const TexCoord& c) {  // bitmap.cpp:82-104
  const float x = c.dudx * c.dudx + c.dvdx * c.dvdx, y = c.dudy * c.dudy + c.dvdy * c.dvdy;
  const float width = sqrtf(fmaxf(x, y));
  const uint32_t ml = t.levels;
  const float level = float(ml - 1) + log2f(fmaxf(width, 1e-8f));
  if (level < 0.f) return tex_bilinear(s, t, 0, c.u, c.v);
  if (level >= float(ml - 1)) return tex_bilinear(s, t, ml - 1, c.u, c.v);
  const uint32_t down = uint32_t(floorf(level));
  const float wgt = level - float(down);
  const f3 lo = tex_bilinear(s, t, down, c.u, c.v), hi = tex_bilinear(s, t, down + 1, c.u, c.v);
  return lo * (1.f - wgt) + hi * wgt;
}

__device__ inline f3 tex_anisotropic(const DevScene& s, const TexRec& t, const TexCoord& c) {  // bitmap.cpp:106-134
  const float x = c.dudx * c.dudx + c.dvdx * c.dvdx, y = c.dudy * c.dudy + c.dvdy * c.dvdy;
  const float big = sqrtf(fmaxf(x, y));
  float small = sqrtf(fminf(x, y));
  const float bx = x < y ? c.dudy : c.dudx, by = x < y ? c.dvdy : c.dvdx;
  float sx = x < y ? c.dudx : c.dudy, sy = x < y ? c.dvdx : c.dvdy;
  if ((small * t.aniso < big) && (small > 0.f)) {
    const float scale = big / (small * t.aniso);
    sx *= scale;
    sy *= scale;
    small *= scale;
  }
  if (small == 0.f) return tex_bilinear(s, t, 0, c.u, c.v);
  const uint32_t ml = t.levels;
  const float level = float(ml - 1) + log2f(fmaxf(small, 1e-8f));
  if (level < 0.f) return tex_ewa(s, t, 0, c.u, c.v, bx, by, sx, sy);
  if (level >= float(ml - 1)) return tex_bilinear(s, t, ml - 1, c.u, c.v);
  const uint32_t down = uint32_t(floorf(level));
  const float wgt = level - float(down);
  const f3 lo = tex_ewa(s, t, down, c.u, c.v, bx, by, sx, sy);
  f3 hi = splat(0.f);
  if (wgt != 0.f) hi = tex_ewa(s, t, down + 1, c.u, c.v, bx, by, sx, sy);
  return lo * (1.f - wgt) + hi * wgt;
}

__device__ inline f3 tex_bitmap(const DevScene& s, const TexRec& t, const TexCoord& c) {  // bitmap.cpp:325-338
  switch (t.filter) {
    case RAD_FILTER_NEAREST:
      return tex_nearest(s, t, c.u, c.v);
    case RAD_FILTER_TRILINEAR:
      return tex_trilinear(s, t, c);
    case RAD_FILTER_ANISOTROPIC:
      return tex_anisotropic(s, t, c);
    default:
      return tex_bilinear(s, t, 0, c.u, c.v);
  }
}

// Texture<T>::Eval (texture.h:57-62); mono textures return their value in .x.
__device__ inline f3 tex_eval(const DevScene& s, uint32_t id, TexCoord c) {
  const TexRec& t = s.textures[id];
  if (t.type != TEX_CHESSBOARD) return tex_bitmap(s, t, c);
  // chessboard.cpp:33-55
  const float* M = t.to_uv;
  const float r0 = M[0] * c.u + M[1] * c.v + M[3], r1 = M[4] * c.u + M[5] * c.v + M[7];
  const float r3 = M[12] * c.u + M[13] * c.v + M[15];
  const float tu = r0 / r3, tv = r1 / r3;
  const bool isRight = tu - floorf(tu) > 0.5f, isUp = tv - floorf(tv) > 0.5f;
  const bool first = isRight == isUp;
  const int32_t nested = first ? t.tex_a : t.tex_b;
  if (nested < 0) return first ? t.color_a : t.color_b;
  if (t.remap) {
    c.u = first ? (tu - 0.5f) * 2.f : tu * 2.f;
    c.v = first ? (tv - 0.5f) * 2.f : tv * 2.f;
  }
  return tex_bitmap(s, s.textures[nested], c);
}

#pragma float_control(pop)

}  // namespace rad
