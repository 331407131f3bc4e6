// Device textures: lookups of `bitmap` (nearest, the reference's bilinear,
// trilinear over the mip chain, EWA) and `chessboard` textures
// (texture/bitmap.cpp:24-247, texture/chessboard.cpp:33-55), the texture
// coordinates of a hit with their screen-space derivatives
// (interaction.cpp:83-111), and the per-hit copy of a textured BSDF record.
//
// Arithmetic is kept uncontracted so that texel selection (the integer
// truncations, the EWA window and its r^2 < 1 test) follows the same float
// values as the serial definition.
#pragma once

#include "shading_dev.h"

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

// Texture coordinates of a surface hit with dP/du and dP/dv
// (mesh_base.cpp:43-66, sphere.cpp:215-249, rectangle.cpp:204-219).
struct HitUV {
  float u, v;
  f3 dpdu, dpdv;
};
__device__ inline HitUV hit_uv(const DevScene& s, const SurfaceHit& h, float bu, float bv, uint32_t prim,
                               uint32_t geom) {
  const ShapeRec& sr = s.shapes[geom];
  HitUV r;
  if (sr.type == SHAPE_MESH) {
    const uint32_t* I = s.idx + 3 * size_t(sr.tri_offset + prim);
    const size_t f0 = sr.vtx_offset + I[0], f1 = sr.vtx_offset + I[1], f2 = sr.vtx_offset + I[2];
    if (sr.flags & SHAPE_HAS_UVS) {
      const f3 p0 = load3(s.pos, f0), p1 = load3(s.pos, f1), p2 = load3(s.pos, f2);
      const float b0 = 1.f - bu - bv;
      const float u0 = s.uv[2 * f0], v0 = s.uv[2 * f0 + 1], u1 = s.uv[2 * f1], v1 = s.uv[2 * f1 + 1];
      const float u2 = s.uv[2 * f2], v2 = s.uv[2 * f2 + 1];
      r.u = u0 * b0 + (u1 * bu + (u2 * bv));
      r.v = v0 * b0 + (v1 * bu + (v2 * bv));
      const float du0 = u1 - u0, dv0 = v1 - v0, du1 = u2 - u0, dv1 = v2 - v0;
      const f3 dp0 = p1 - p0, dp1 = p2 - p0;
      const float det = du0 * dv1 - (dv0 * du1);
      const float invDet = det == 0.f ? 0.f : 1.f / det;
      if (invDet == 0.f) {
        r.dpdu = r.dpdv = splat(0.f);
      } else {
        r.dpdu = (dp0 * dv1 - (dp1 * dv0)) * invDet;
        r.dpdv = (dp0 * (-du1) + (dp1 * du0)) * invDet;
      }
    } else {
      r.u = bu;
      r.v = bv;
      coordinate_system(h.n, r.dpdu, r.dpdv);
    }
  } else if (sr.type == SHAPE_SPHERE) {
    const f3 local = sr.to_local.point(h.p);
    const float rd2 = local.x * local.x + local.y * local.y;
    // Frame-free polar angle (math_ext.h UnitAngleZ)
    const float half = asinf(0.5f * sqrtf(sqnorm(mk(local.x, local.y, local.z - mulsign(1.f, local.z))))) * 2.f;
    const float theta = local.z >= 0.f ? half : kPi - half;
    float phi = atan2f(local.y, local.x);
    if (phi < 0.f) phi += 2.f * kPi;
    r.u = phi / (2.f * kPi);
    r.v = theta / kPi;
    const float rr = sqrtf(rd2), inv = 1.f / rr, cosPhi = local.x * inv, sinPhi = local.y * inv;
    const f3 dv = rr == 0.f ? mk(1.f, 0.f, 0.f) : mk(local.z * cosPhi, local.z * sinPhi, -rr);
    r.dpdu = sr.to_world.vec(mk(-local.y, local.x, 0.f) * (2.f * kPi));
    r.dpdv = sr.to_world.vec(dv * kPi);
  } else {  // the hit record carries the rectangle's local x, y
    r.u = bu * 0.5f + 0.5f;
    r.v = bv * 0.5f + 0.5f;
    r.dpdu = sr.frame_s;
    r.dpdv = sr.frame_t;
  }
  return r;
}

// SurfaceInteraction::BSDF(const RayDifferential&) (interaction.cpp:83-111)
// for a camera ray through screen position (sx, sy).
__device__ inline void uv_footprint(const CameraRec& c, float sx, float sy, const SurfaceHit& h, const HitUV& uv,
                                    TexCoord& tc) {
  const f3 ndc = mk(sx * c.rcp_w, sy * c.rcp_h, 0.f);
  const f3 nearP = c.clip_to_camera.point_h(ndc);
  const f3 o = c.camera_to_world.point(mk(0.f, 0.f, 0.f));
  const f3 Dx = c.camera_to_world.vec(normalize(nearP + c.dx));
  const f3 Dy = c.camera_to_world.vec(normalize(nearP + c.dy));
  const float d = dot(h.n, h.p);
  const float tx = (d - dot(h.n, o)) / dot(h.n, Dx);
  const float ty = (d - dot(h.n, o)) / dot(h.n, Dy);
  const f3 dpdx = (Dx * tx + o) - h.p;
  const f3 dpdy = (Dy * ty + o) - h.p;
  const float a00 = dot(uv.dpdu, uv.dpdu), a01 = dot(uv.dpdu, uv.dpdv), a11 = dot(uv.dpdv, uv.dpdv);
  float invDet = 1.f / fmaf(a00, a11, -(a01 * a01));
  const float b0x = dot(uv.dpdu, dpdx), b1x = dot(uv.dpdv, dpdx), b0y = dot(uv.dpdu, dpdy), b1y = dot(uv.dpdv, dpdy);
  invDet = isfinite(invDet) ? invDet : 0.f;
  tc.dudx = fmaf(a11, b0x, -(a01 * b1x)) * invDet;
  tc.dvdx = fmaf(a00, b1x, -(a01 * b0x)) * invDet;
  tc.dudy = fmaf(a11, b0y, -(a01 * b1y)) * invDet;
  tc.dvdy = fmaf(a00, b1y, -(a01 * b0y)) * invDet;
}

// Texture coordinates of a hit from its (t, u, v, prim) record; camera: the
// ray is the camera ray through screen position (sx, sy), whose footprint
// gives the derivatives (zero otherwise).
__device__ inline TexCoord hit_texcoord(const DevScene& s, const SurfaceHit& h, float4 hh, uint32_t geom, bool camera,
                                        float sx, float sy) {
  const HitUV uv = hit_uv(s, h, hh.y, hh.z, __float_as_uint(hh.w), geom);
  TexCoord tc{uv.u, uv.v, 0.f, 0.f, 0.f, 0.f};
  if (camera) uv_footprint(s.cam, sx, sy, h, uv, tc);
  return tc;
}

// A BSDF record with its textured slots evaluated at the hit (the
// reference evaluates each Texture member inside Eval / Sample at the same
// interaction).
__device__ inline void resolve_textures(const DevScene& s, BsdfRec& b, const TexCoord& tc) {
  uint32_t m = b.tex_mask;
  while (m) {
    const int k = __ffs(m) - 1;
    m &= m - 1;
    const f3 v = tex_eval(s, b.tex[k], tc);
    switch (k) {
      case RAD_TEX_REFLECTANCE: b.reflectance = v; break;
      case RAD_TEX_TRANSMITTANCE: b.transmittance = v; break;
      case RAD_TEX_ETA: b.eta3 = v; break;
      case RAD_TEX_K: b.k3 = v; break;
      case RAD_TEX_ALPHA_U: b.alpha_u = v.x; break;
      case RAD_TEX_ALPHA_V: b.alpha_v = v.x; break;
      case RAD_TEX_BASE_COLOR: b.base_color = v; break;
      case RAD_TEX_ROUGHNESS: b.roughness = v.x; break;
      case RAD_TEX_ANISOTROPIC: b.anisotropic = v.x; break;
      case RAD_TEX_SPEC_TRANS: b.spec_trans = v.x; break;
      case RAD_TEX_SHEEN: b.sheen = v.x; break;
      case RAD_TEX_SHEEN_TINT: b.sheen_tint = v.x; break;
      case RAD_TEX_FLATNESS: b.flatness = v.x; break;
      case RAD_TEX_SPEC_TINT: b.spec_tint = v.x; break;
      case RAD_TEX_METALLIC: b.metallic = v.x; break;
      case RAD_TEX_CLEARCOAT: b.clearcoat = v.x; break;
      case RAD_TEX_CLEARCOAT_GLOSS: b.clearcoat_gloss = v.x; break;
      case RAD_TEX_DIFFUSE: b.diffuse3 = v; break;
      case RAD_TEX_SPECULAR: b.spec3 = v; break;
      case RAD_TEX_OPACITY: b.opacity = fminf(fmaxf(v.x, 0.f), 1.f); break;
      default: break;
    }
  }
}

// The wrapper chain of `top` (at most two wrappers and the wrapped BSDF)
// copied into chain[0..2] with textures evaluated; chain[k].nested is
// re-pointed at chain[k + 1], so bsdf_eval_w / bsdf_sample_w can walk it.
__device__ inline void resolve_chain(const DevScene& s, const BsdfRec& top, const TexCoord& tc, BsdfRec (&chain)[3]) {
  chain[0] = top;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    resolve_textures(s, chain[k], tc);
    const bool wrap = chain[k].type == BSDF_TWO_SIDE || chain[k].type == BSDF_MASK;
    if (k < 2 && wrap) {
      chain[k + 1] = s.bsdfs[chain[k].nested];
      chain[k].nested = k + 1;
    } else {
      break;
    }
  }
}

#pragma float_control(pop)

}  // namespace rad
