// Texture coordinates of a hit with their screen-space derivatives
// (interaction.cpp:83-111), and the per-hit copy of a textured BSDF record
// (lookups: texture_eval.h).
#pragma once

#include "shading_dev.h"
#include "texture_eval.h"

namespace rad {

#pragma float_control(push)
#pragma clang fp contract(off)

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
    const float4* T = s.tris + 6 * size_t(sr.tri_offset + prim);  // DevScene::tris
    if (sr.flags & SHAPE_HAS_UVS) {
      const float4 t0 = T[0], t1 = T[1], t2 = T[2], t4 = T[4], t5 = T[5];
      const f3 p0 = mk(t0.x, t0.y, t0.z), p1 = mk(t0.w, t1.x, t1.y), p2 = mk(t1.z, t1.w, t2.x);
      const float b0 = 1.f - bu - bv;
      const float u0 = t4.z, v0 = t4.w, u1 = t5.x, v1 = t5.y;
      const float u2 = t5.z, v2 = t5.w;
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
