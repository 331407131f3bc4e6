// ORACLE / TEST INFRASTRUCTURE ONLY (see ref_math.h).
// Textures of the restatement: the constant texture (render/texture.h:42-89),
// `bitmap` with its mip chain and four filters (texture/bitmap.cpp:24-247),
// `chessboard` (texture/chessboard.cpp:14-62), the texture-space footprint of
// a camera ray (interaction.cpp:83-111) and the per-hit parameter lookup of
// the BSDFs (every `_x->Eval(si)` of bsdf/*.cpp).
#include "ref_scene.h"

namespace ref {

namespace {

V3 Scaled(const V3& a, Float s) { return a * s; }

// MipMap::Wrap (bitmap.cpp:136-156), float and integer forms.
Float WrapF(Float v, int wrap) {
  if (wrap == RAD_WRAP_REPEAT) return std::clamp(v - std::floor(v), Float(0), Float(1));
  return std::clamp(v, Float(0), Float(1));
}

}  // namespace

const V3& TexImage::At(uint32_t x, uint32_t y) const { return px[size_t(y) * w + x]; }

// bitmap.cpp:158-166: integer lookup used to build the mip chain.
V3 Texture::BilinearAt(const TexImage& img, uint32_t u, uint32_t v) const {
  uint32_t pu, pv;
  if (wrap == RAD_WRAP_REPEAT) {
    pu = u % img.w;
    pv = v % img.h;
  } else {
    pu = std::clamp(u, 0u, img.w - 1);
    pv = std::clamp(v, 0u, img.h - 1);
  }
  Float fu = Float(pu) + Float(1);
  Float fv = Float(pv) + Float(1);
  return Bilinear4(img, pu, pv, fu, fv);
}

// bitmap.cpp:192-206.  The neighbour index is formed in unsigned arithmetic
// and then clamped, so below texel 0 it becomes the last texel of the row.
V3 Texture::Bilinear4(const TexImage& img, uint32_t pu, uint32_t pv, Float fu, Float fv) const {
  int32_t dpu = (fu > Float(pu) + 0.5f) ? 1 : -1;
  int32_t dpv = (fv > Float(pv) + 0.5f) ? 1 : -1;
  uint32_t apu = std::clamp(uint32_t(pu + uint32_t(dpu)), 0u, img.w - 1);
  uint32_t apv = std::clamp(uint32_t(pv + uint32_t(dpv)), 0u, img.h - 1);
  Float du = std::min(std::abs(fu - Float(pu) - Float(0.5)), Float(1));
  Float dv = std::min(std::abs(fv - Float(pv) - Float(0.5)), Float(1));
  const V3& u0v0 = img.At(pu, pv);
  const V3& u1v0 = img.At(apu, pv);
  const V3& u0v1 = img.At(pu, apv);
  const V3& u1v1 = img.At(apu, apv);
  return (Scaled(u0v0, 1 - du) + Scaled(u1v0, du)) * (1 - dv) + (Scaled(u0v1, 1 - du) + Scaled(u1v1, du)) * dv;
}

// bitmap.cpp:168-178
V3 Texture::Nearest(const TexImage& img, const V2& uv) const {
  Float u = WrapF(uv.x, wrap), v = WrapF(uv.y, wrap);
  Float fu = u * Float(img.w - 1);
  Float fv = v * Float(img.h - 1);
  return img.At(uint32_t(fu), uint32_t(fv));
}

// bitmap.cpp:180-190
V3 Texture::Bilinear(const TexImage& img, const V2& uv) const {
  Float u = WrapF(uv.x, wrap), v = WrapF(uv.y, wrap);
  Float fu = u * Float(img.w - 1);
  Float fv = v * Float(img.h - 1);
  return Bilinear4(img, uint32_t(fu), uint32_t(fv), fu, fv);
}

// bitmap.cpp:208-212: every level is read bilinearly whatever the filter.
V3 Texture::Isotropic(const V2& uv, size_t level) const {
  size_t l = std::clamp(level, size_t(0), mips.size() - 1);
  return Bilinear(mips[l], uv);
}

// bitmap.cpp:214-246.  The weights are the level's texel grid, but the
// lookups read level 0 by nearest neighbour (EvalNearest), as the reference does.
V3 Texture::Ewa(const V2& uv, size_t level, const V2& d0, const V2& d1) const {
  const TexImage& map = mips[level];
  V2 st(uv.x * Float(map.w), uv.y * Float(map.h));
  V2 dst0(d0.x * Float(map.w), d0.y * Float(map.h));
  V2 dst1(d1.x * Float(map.w), d1.y * Float(map.h));
  Float A = dst0.y * dst0.y + dst1.y * dst1.y + 1;
  Float B = -2 * (dst0.x * dst0.y + dst1.x * dst1.y);
  Float C = dst0.x * dst0.x + dst1.x * dst1.x + 1;
  Float invF = 1 / (A * C - B * B * Float(0.25));
  A *= invF;
  B *= invF;
  C *= invF;
  Float det = -B * B + 4 * A * C;
  Float invDet = 1 / det;
  Float uSqrt = std::sqrt(det * C), vSqrt = std::sqrt(det * A);
  int32_t s0 = int32_t(std::ceil(st.x - 2 * invDet * uSqrt));
  int32_t s1 = int32_t(std::floor(st.x + 2 * invDet * uSqrt));
  int32_t t0 = int32_t(std::ceil(st.y - 2 * invDet * vSqrt));
  int32_t t1 = int32_t(std::floor(st.y + 2 * invDet * vSqrt));
  V2 invR(1 / Float(map.w), 1 / Float(map.h));
  V3 sum(0);
  Float sumWts = 0;
  for (int32_t it = t0; it <= t1; ++it) {
    Float tt = Float(it) - st.y;
    for (int32_t is = s0; is <= s1; ++is) {
      Float ss = Float(is) - st.x;
      Float r2 = A * ss * ss + B * ss * tt + C * tt * tt;
      if (r2 < 1) {
        Float weight = std::exp(Float(-2) * r2) - std::exp(Float(-2));
        V3 value = Nearest(mips[0], V2(Float(is) * invR.x, Float(it) * invR.y));
        sum += value * weight;
        sumWts += weight;
      }
    }
  }
  return sum / sumWts;
}

// bitmap.cpp:82-104
V3 Texture::Trilinear(const V2& uv, const V2& dx, const V2& dy) const {
  Float x = dx.x * dx.x + dx.y * dx.y;
  Float y = dy.x * dy.x + dy.y * dy.y;
  Float width = std::sqrt(std::max(x, y));
  const size_t ml = mips.size();
  Float level = Float(ml - 1) + std::log2(std::max(width, Float(1e-8)));
  if (level < 0) return Isotropic(uv, 0);
  if (level >= Float(ml - 1)) return Isotropic(uv, ml - 1);
  size_t down = size_t(std::floor(level));
  Float weight = level - Float(down);
  V3 low = Isotropic(uv, down);
  V3 high = Isotropic(uv, down + 1);
  return low * (1 - weight) + high * weight;
}

// bitmap.cpp:106-134
V3 Texture::Anisotropic(const V2& uv, const V2& dx, const V2& dy) const {
  Float x = dx.x * dx.x + dx.y * dx.y;
  Float y = dy.x * dy.x + dy.y * dy.y;
  Float big = std::sqrt(std::max(x, y));
  Float small = std::sqrt(std::min(x, y));
  V2 bigduv = x < y ? dy : dx;
  V2 smallduv = x < y ? dx : dy;
  if ((small * aniso < big) && (small > 0)) {
    Float scale = big / (small * aniso);
    smallduv = V2(smallduv.x * scale, smallduv.y * scale);
    small *= scale;
  }
  if (small == 0) return Isotropic(uv, 0);
  const size_t ml = mips.size();
  Float level = Float(ml - 1) + std::log2(std::max(small, Float(1e-8)));
  if (level < 0) return Ewa(uv, 0, bigduv, smallduv);
  if (level >= Float(ml - 1)) return Isotropic(uv, ml - 1);
  size_t down = size_t(std::floor(level));
  Float weight = level - Float(down);
  V3 low = Ewa(uv, down, bigduv, smallduv);
  V3 high(0);
  if (weight != 0) high = Ewa(uv, down + 1, bigduv, smallduv);
  return low * (1 - weight) + high * weight;
}

// bitmap.cpp:27-61
void Texture::BuildMips(TexImage level0, int32_t maxLevel) {
  if (maxLevel != 0 && ((level0.w % 2 != 0) || (level0.h % 2 != 0))) maxLevel = 0;
  uint32_t realMax = maxLevel < 0 ? 1 + uint32_t(std::log2(double(std::max(level0.w, level0.h)))) : uint32_t(maxLevel);
  realMax = std::max(realMax, 1u);
  mips.clear();
  mips.push_back(std::move(level0));
  for (uint32_t a = 1; a < realMax; a++) {
    const TexImage& last = mips[a - 1];
    TexImage map;
    map.w = std::max(1u, last.w / 2);
    map.h = std::max(1u, last.h / 2);
    map.px.resize(size_t(map.w) * map.h);
    for (uint32_t j = 0; j < map.h; j++) {
      for (uint32_t i = 0; i < map.w; i++) {
        V3 u0v0 = BilinearAt(last, i * 2, j * 2);
        V3 u1v0 = BilinearAt(last, i * 2 + 1, j * 2);
        V3 u0v1 = BilinearAt(last, i * 2, j * 2 + 1);
        V3 u1v1 = BilinearAt(last, i * 2 + 1, j * 2 + 1);
        map.px[size_t(j) * map.w + i] = (u0v0 + u1v0 + u0v1 + u1v1) * Float(0.25);
      }
    }
    mips.push_back(std::move(map));
  }
}

V3 Texture::Eval(const V2& uv, const V2& dx, const V2& dy) const {
  if (type == RAD_TEXTURE_CHESSBOARD) {  // chessboard.cpp:33-55
    V3 t = toUV.ApplyAffineToWorld(V3(uv.x, uv.y, 0));
    V2 tuv(t.x, t.y);
    V2 ab(tuv.x - std::floor(tuv.x), tuv.y - std::floor(tuv.y));
    bool isRight = ab.x > 0.5f;
    bool isUp = ab.y > 0.5f;
    if (isRight == isUp) {
      V2 nuv = remap ? V2((tuv.x - Float(0.5)) * 2, (tuv.y - Float(0.5)) * 2) : uv;
      return a ? a->Eval(nuv, dx, dy) : colorA;
    }
    V2 nuv = remap ? V2(tuv.x * 2, tuv.y * 2) : uv;
    return b ? b->Eval(nuv, dx, dy) : colorB;
  }
  switch (filter) {  // bitmap.cpp:325-338
    case RAD_FILTER_NEAREST:
      return Nearest(mips[0], uv);
    case RAD_FILTER_BILINEAR:
      return Bilinear(mips[0], uv);
    case RAD_FILTER_TRILINEAR:
      return Trilinear(uv, dx, dy);
    case RAD_FILTER_ANISOTROPIC:
      return Anisotropic(uv, dx, dy);
    default:
      return V3(0);
  }
}

// Texture<T>::Read of level 0 (texture.h:67-72); a chessboard reads 0 and
// counts as one texel (chessboard.cpp:29-30, 57-59).
double Texture::MeanLuminance() const {
  if (type == RAD_TEXTURE_CHESSBOARD) return 0.0;
  const TexImage& img = mips[0];
  double sum = 0;
  for (uint32_t y = 0; y < img.h; y++)
    for (uint32_t x = 0; x < img.w; x++) sum += (channels == 3 ? Luminance(img.At(x, y)) : img.At(x, y).x);
  return sum / (double(img.h) * double(img.w));
}

std::unique_ptr<Texture> MakeTextureBitmap(const rad_texture_desc& d) {
  if (d.type != RAD_TEXTURE_BITMAP) throw std::runtime_error("texture: not a bitmap");
  if (d.width == 0 || d.height == 0 || !d.texels) throw std::runtime_error("bitmap: empty image");
  if (d.channels != 1 && d.channels != 3) throw std::runtime_error("bitmap: channels must be 1 or 3");
  auto t = std::make_unique<Texture>();
  t->type = RAD_TEXTURE_BITMAP;
  t->channels = d.channels;
  t->filter = d.filter;
  t->wrap = d.wrap;
  t->aniso = d.anisotropic_level;
  TexImage img;
  img.w = d.width;
  img.h = d.height;
  img.px.resize(size_t(d.width) * d.height);
  for (size_t i = 0; i < img.px.size(); i++) {
    const float* p = d.texels + i * size_t(d.channels);
    img.px[i] = d.channels == 3 ? V3(p[0], p[1], p[2]) : V3(p[0], 0, 0);
  }
  t->BuildMips(std::move(img), d.max_level);
  return t;
}

void BuildTextures(const rad_texture_desc* descs, uint32_t count, std::vector<std::unique_ptr<Texture>>& out) {
  out.clear();
  out.resize(count);
  for (uint32_t i = 0; i < count; i++)
    if (descs[i].type == RAD_TEXTURE_BITMAP) out[i] = MakeTextureBitmap(descs[i]);
  for (uint32_t i = 0; i < count; i++) {
    const rad_texture_desc& d = descs[i];
    if (d.type == RAD_TEXTURE_BITMAP) continue;
    if (d.type != RAD_TEXTURE_CHESSBOARD) throw std::runtime_error("texture: unknown type");
    auto t = std::make_unique<Texture>();
    t->type = RAD_TEXTURE_CHESSBOARD;
    t->channels = d.channels;
    t->toUV = Transform(Mat4::FromRowMajor(d.to_uv));
    t->remap = d.remap != 0;
    t->colorA = d.channels == 3 ? V3(d.color_a[0], d.color_a[1], d.color_a[2]) : V3(d.color_a[0], 0, 0);
    t->colorB = d.channels == 3 ? V3(d.color_b[0], d.color_b[1], d.color_b[2]) : V3(d.color_b[0], 0, 0);
    auto nested = [&](int32_t k) -> const Texture* {
      if (k == 0) return nullptr;
      if (k < 1 || uint32_t(k) > count || descs[k - 1].type != RAD_TEXTURE_BITMAP)
        throw std::runtime_error("chessboard: color_a / color_b must be constants or bitmaps");
      return out[k - 1].get();
    };
    t->a = nested(d.tex_a);
    t->b = nested(d.tex_b);
    out[i] = std::move(t);
  }
}

// SurfaceInteraction::BSDF(const RayDifferential&) (interaction.cpp:83-111).
void ComputeUVDifferentials(SurfaceInteraction& si, const RayDiff& rd) {
  if (!rd.has) return;
  Float d = Dot(si.N, si.P);
  Float tx = (d - Dot(si.N, rd.Ox)) / Dot(si.N, rd.Dx);
  Float ty = (d - Dot(si.N, rd.Oy)) / Dot(si.N, rd.Dy);
  V3 dpdx = (rd.Dx * tx + rd.Ox) - si.P;
  V3 dpdy = (rd.Dy * ty + rd.Oy) - si.P;
  Float a00 = Dot(si.dPdU, si.dPdU), a01 = Dot(si.dPdU, si.dPdV), a11 = Dot(si.dPdV, si.dPdV);
  Float invDet = Rcp(Fmadd(a00, a11, -(a01 * a01)));
  Float b0x = Dot(si.dPdU, dpdx), b1x = Dot(si.dPdV, dpdx), b0y = Dot(si.dPdU, dpdy), b1y = Dot(si.dPdV, dpdy);
  invDet = std::isfinite(invDet) ? invDet : 0;
  si.dUVdX = V2(Fmadd(a11, b0x, -(a01 * b1x)) * invDet, Fmadd(a00, b1x, -(a01 * b0x)) * invDet);
  si.dUVdY = V2(Fmadd(a11, b0y, -(a01 * b1y)) * invDet, Fmadd(a00, b1y, -(a01 * b0y)) * invDet);
}

// The BSDF of a hit with its textured parameters evaluated at the hit: the
// reference evaluates each Texture member inside Eval / Sample / Pdf at the
// same interaction, which is the same as evaluating them once per hit.
const Bsdf* Scene::HitBsdf(const SurfaceInteraction& si, std::unique_ptr<Bsdf>& holder) const {
  const int32_t idx = si.shape->bsdfIndex;
  if (idx < 0 || !bsdfTextured[size_t(idx)]) return si.shape->bsdf;
  std::vector<rad_bsdf_desc> local = bsdfDescs;
  for (int32_t j = idx, guard = 0; j >= 0 && guard < 3; guard++) {
    rad_bsdf_desc& d = local[size_t(j)];
    for (int s = 0; s < RAD_TEX_SLOTS; s++) {
      if (d.tex[s] == 0) continue;
      V3 v = textures[size_t(d.tex[s] - 1)]->Eval(si.UV, si.dUVdX, si.dUVdY);
      float* f3p = nullptr;
      float* f1p = nullptr;
      switch (s) {
        case RAD_TEX_REFLECTANCE: f3p = d.reflectance; break;
        case RAD_TEX_TRANSMITTANCE: f3p = d.transmittance; break;
        case RAD_TEX_ETA: f3p = d.eta; break;
        case RAD_TEX_K: f3p = d.k; break;
        case RAD_TEX_BASE_COLOR: f3p = d.base_color; break;
        case RAD_TEX_DIFFUSE: f3p = d.diffuse; break;
        case RAD_TEX_SPECULAR: f3p = d.specular_reflectance; break;
        case RAD_TEX_ALPHA_U: f1p = &d.alpha_u; break;
        case RAD_TEX_ALPHA_V: f1p = &d.alpha_v; break;
        case RAD_TEX_ROUGHNESS: f1p = &d.roughness; break;
        case RAD_TEX_ANISOTROPIC: f1p = &d.anisotropic; break;
        case RAD_TEX_SPEC_TRANS: f1p = &d.spec_trans; break;
        case RAD_TEX_SHEEN: f1p = &d.sheen; break;
        case RAD_TEX_SHEEN_TINT: f1p = &d.sheen_tint; break;
        case RAD_TEX_FLATNESS: f1p = &d.flatness; break;
        case RAD_TEX_SPEC_TINT: f1p = &d.spec_tint; break;
        case RAD_TEX_METALLIC: f1p = &d.metallic; break;
        case RAD_TEX_CLEARCOAT: f1p = &d.clearcoat; break;
        case RAD_TEX_CLEARCOAT_GLOSS: f1p = &d.clearcoat_gloss; break;
        case RAD_TEX_OPACITY: f1p = &d.opacity; break;
        default: break;
      }
      if (f3p) f3p[0] = v.x, f3p[1] = v.y, f3p[2] = v.z;
      if (f1p) *f1p = v.x;
    }
    j = (d.type == RAD_BSDF_TWO_SIDE || d.type == RAD_BSDF_MASK) ? d.nested : -1;
  }
  holder = MakeBsdf(local.data(), uint32_t(local.size()), uint32_t(idx), quirks, 0, specWeights.data());
  return holder.get();
}

}  // namespace ref
