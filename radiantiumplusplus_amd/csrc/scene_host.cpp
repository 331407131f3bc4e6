#include "scene_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace rad {

std::string HipErr(hipError_t e) { return std::string(hipGetErrorName(e)) + ": " + hipGetErrorString(e); }

void HipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + " failed: " + HipErr(e));
}

void DeviceBuffer::reset() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

void DeviceBuffer::alloc(size_t n) {
  if (n <= bytes && ptr) return;
  reset();
  if (n == 0) return;
  HipCheck(hipMalloc(&ptr, n), "hipMalloc");
  bytes = n;
}

void DeviceBuffer::upload(const void* src, size_t n, hipStream_t s) {
  alloc(std::max<size_t>(n, 16));
  if (n) HipCheck(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
}

namespace {

// Row-major 4x4 with a float cofactor inverse (Transform, transform.cpp:5-16).
struct M4 {
  float m[16];
  static M4 Identity() {
    M4 r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.f;
    return r;
  }
  static M4 From(const float* a) {
    M4 r;
    std::memcpy(r.m, a, sizeof(r.m));
    return r;
  }
  M4 operator*(const M4& o) const {
    M4 r{};
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        float s = 0;
        for (int k = 0; k < 4; k++) s += m[i * 4 + k] * o.m[k * 4 + j];
        r.m[i * 4 + j] = s;
      }
    return r;
  }
  bool IsIdentity() const {
    for (int i = 0; i < 16; i++)
      if (m[i] != ((i % 5) == 0 ? 1.f : 0.f)) return false;
    return true;
  }
  M4 Inverse() const {
    // Gauss-Jordan in double, rounded to float
    double a[4][8];
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        a[i][j] = m[i * 4 + j];
        a[i][j + 4] = i == j ? 1.0 : 0.0;
      }
    for (int c = 0; c < 4; c++) {
      int p = c;
      for (int r = c + 1; r < 4; r++)
        if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
      if (std::fabs(a[p][c]) < 1e-300) throw std::runtime_error("transform matrix is not invertible");
      if (p != c)
        for (int j = 0; j < 8; j++) std::swap(a[p][j], a[c][j]);
      double inv = 1.0 / a[c][c];
      for (int j = 0; j < 8; j++) a[c][j] *= inv;
      for (int r = 0; r < 4; r++) {
        if (r == c) continue;
        double f = a[r][c];
        if (f == 0) continue;
        for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
      }
    }
    M4 r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) r.m[i * 4 + j] = float(a[i][j + 4]);
    return r;
  }
  f3 Point(f3 v) const {  // with homogeneous divide
    float r0 = m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3];
    float r1 = m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7];
    float r2 = m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11];
    float r3 = m[12] * v.x + m[13] * v.y + m[14] * v.z + m[15];
    return mk(r0 / r3, r1 / r3, r2 / r3);
  }
  f3 Vec(f3 v) const {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z);
  }
  f3 TVec(f3 v) const {  // transpose of the linear part
    return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
              m[2] * v.x + m[6] * v.y + m[10] * v.z);
  }
  Affine ToAffine() const {
    Affine a;
    for (int i = 0; i < 12; i++) a.m[i] = m[i];
    return a;
  }
};

const float kCubeV[24][3] = {{1, -1, -1}, {1, -1, 1},   {-1, -1, 1}, {-1, -1, -1}, {1, 1, -1},  {-1, 1, -1},
                             {-1, 1, 1},  {1, 1, 1},    {1, -1, -1}, {1, 1, -1},   {1, 1, 1},   {1, -1, 1},
                             {1, -1, 1},  {1, 1, 1},    {-1, 1, 1},  {-1, -1, 1},  {-1, -1, 1}, {-1, 1, 1},
                             {-1, 1, -1}, {-1, -1, -1}, {1, 1, -1},  {1, -1, -1},  {-1, -1, -1}, {-1, 1, -1}};
const float kCubeN[6][3] = {{0, -1, 0}, {0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {-1, 0, 0}, {0, 0, -1}};
const uint32_t kCubeT[12][3] = {{0, 1, 2},    {3, 0, 2},    {4, 5, 6},    {7, 4, 6},    {8, 9, 10},   {11, 8, 10},
                                {12, 13, 14}, {15, 12, 14}, {16, 17, 18}, {19, 16, 18}, {20, 21, 22}, {23, 20, 22}};

float4 F4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
float AsF(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

void GrowBounds(float* b, f3 p) {
  b[0] = std::min(b[0], p.x);
  b[1] = std::min(b[1], p.y);
  b[2] = std::min(b[2], p.z);
  b[3] = std::max(b[3], p.x);
  b[4] = std::max(b[4], p.y);
  b[5] = std::max(b[5], p.z);
}

}  // namespace

// Fresnel::DiffuseReflectance (fresnel.cpp:79-84), Horner-evaluated with
// fused multiply-adds as math_base.h:90-105 does.
static float DiffuseReflectance(float eta) {
  float invEta = 1.0f / eta;
  if (eta < 1.f) return std::fma(0.0636f, invEta, std::fma(eta, std::fma(eta, -1.4399f, 0.7099f), 0.6681f));
  const float c[6] = {0.919317f, -3.4793f, 6.75335f, -7.80989f, 4.98554f, -1.36881f};
  float acc = c[5];
  for (int i = 4; i >= 0; i--) acc = std::fma(invEta, acc, c[i]);
  return acc;
}

static float Lum(const float* c) { return (c[0] * 0.212671f) + (c[1] * 0.715160f) + (c[2] * 0.072169f); }

BsdfRec MakeBsdfRec(const rad_bsdf_desc& d) {
  BsdfRec b{};
  auto v3 = [](const float* a) { return mk(a[0], a[1], a[2]); };
  b.type = d.type;
  b.beckmann = d.distribution == RAD_MICROFACET_BECKMANN;
  b.sample_visible = d.sample_visible;
  b.disney_flags = d.disney_flags;
  b.reflectance = v3(d.reflectance);
  b.transmittance = v3(d.transmittance);
  b.eta3 = v3(d.eta);
  b.k3 = v3(d.k);
  b.alpha_u = d.alpha_u;
  b.alpha_v = d.alpha_v;
  b.base_color = v3(d.base_color);
  b.roughness = d.roughness;
  b.anisotropic = d.anisotropic;
  b.spec_trans = d.spec_trans;
  b.sheen = d.sheen;
  b.sheen_tint = d.sheen_tint;
  b.flatness = d.flatness;
  b.spec_tint = d.spec_tint;
  b.metallic = d.metallic;
  b.clearcoat = d.clearcoat;
  b.clearcoat_gloss = d.clearcoat_gloss;
  b.spec_rate = d.specular_sampling_rate;
  b.cc_rate = d.clearcoat_sampling_rate;
  b.diff_rate = d.diffuse_sampling_rate;
  switch (d.type) {
    case RAD_BSDF_DIFFUSE:
      b.lobes = LOBE_DIFFUSE | LOBE_REFLECTION;  // diffuse.cpp:42
      break;
    case RAD_BSDF_ROUGH_METAL:
      b.lobes = LOBE_GLOSSY | LOBE_REFLECTION;  // rough_metal.cpp:23
      break;
    case RAD_BSDF_ROUGH_GLASS:
      b.lobes = LOBE_GLOSSY | LOBE_REFLECTION | LOBE_TRANSMISSION;  // rough_glass.cpp:224
      b.eta = d.int_ior / d.ext_ior;
      b.inv_eta = d.ext_ior / d.int_ior;
      break;
    case RAD_BSDF_DISNEY: {  // disney.cpp:86-109
      bool hasSpecTrans = d.disney_flags & RAD_DISNEY_HAS_SPEC_TRANS;
      if (d.disney_flags & RAD_DISNEY_HAS_ETA) {
        b.eta = d.disney_eta;
        if (hasSpecTrans && b.eta == 1.f) b.eta = 1.001f;
      } else {
        float specular = d.specular;
        if (hasSpecTrans && specular == 0.f) specular = 0.001f;
        b.eta = 2.0f * (1.0f / (1.0f - std::sqrt(0.08f * specular))) - 1.0f;
      }
      b.inv_eta = 1.0f / b.eta;
      b.lobes = LOBE_DIFFUSE | LOBE_REFLECTION;
      if (d.disney_flags & RAD_DISNEY_HAS_CLEARCOAT) b.lobes |= LOBE_GLOSSY;
      if (hasSpecTrans) b.lobes |= LOBE_TRANSMISSION;
      break;
    }
    case RAD_BSDF_PLASTIC:
    case RAD_BSDF_ROUGH_PLASTIC: {  // plastic.cpp:20-40, rough_plastic.cpp:20-55
      b.lobes = LOBE_DIFFUSE | LOBE_REFLECTION | LOBE_DELTA;
      b.eta = d.int_ior / d.ext_ior;
      b.inv_eta = 1.0f / b.eta;
      b.diffuse3 = v3(d.diffuse);
      b.spec3 = v3(d.specular_reflectance);
      b.nonlinear = d.nonlinear != 0;
      b.fdr_int = DiffuseReflectance(1.0f / b.eta);
      b.inv_eta2 = 1.0f / (b.eta * b.eta);
      double dm = Lum(d.diffuse), sm = Lum(d.specular_reflectance);
      b.spec_weight = float(sm / (dm + sm));
      b.alpha_v = b.alpha_u;  // isotropic "alpha"
      break;
    }
    case RAD_BSDF_PERFECT_MIRROR:
      b.lobes = LOBE_DELTA | LOBE_REFLECTION;  // perfect_mirror.cpp:18
      break;
    case RAD_BSDF_PERFECT_GLASS:
      b.lobes = LOBE_DELTA | LOBE_REFLECTION | LOBE_TRANSMISSION;  // perfect_glass.cpp:17
      b.eta = d.int_ior / d.ext_ior;
      b.inv_eta = d.ext_ior / d.int_ior;
      break;
    case RAD_BSDF_TWO_SIDE:
    case RAD_BSDF_MASK:  // lobes: the wrapped BSDF's, filled in once the array is complete
      b.nested = d.nested;
      b.opacity = std::min(std::max(d.opacity, 0.0f), 1.0f);
      break;
    default:
      throw std::runtime_error("unsupported BSDF type " + std::to_string(d.type));
  }
  return b;
}

// ---------------------------------------------------------------- textures
// A bitmap's mip chain (texture/bitmap.cpp:27-61): levels of half size (at
// least 1), each texel the mean of four bilinear lookups at integer
// coordinates 2i, 2i + 1 of the level above; with those coordinates the
// reference's bilinear weights are exactly 1/2 towards the next texel
// (clamped at the last one), after the integer wrap of the coordinates.
// Odd sizes at level 0 disable the chain; max_level < 0 asks for
// 1 + floor(log2(max(w, h))) levels.
static void BuildMipChain(const rad_texture_desc& d, TexRec& t, std::vector<float>& pool) {
  const uint32_t C = uint32_t(d.channels);
  int32_t maxLevel = d.max_level;
  if (maxLevel != 0 && ((d.width % 2 != 0) || (d.height % 2 != 0))) maxLevel = 0;
  uint32_t levels = maxLevel < 0 ? 1 + uint32_t(std::log2(double(std::max(d.width, d.height)))) : uint32_t(maxLevel);
  levels = std::max(levels, 1u);
  if (levels > uint32_t(kMaxMips)) throw std::runtime_error("bitmap: more than 16 mip levels");
  t.levels = levels;
  if (pool.size() + size_t(d.width) * d.height * C * 2 > 0xffffffffull) throw std::runtime_error("textures exceed 4G texels");
  t.off[0] = uint32_t(pool.size());
  t.w[0] = d.width;
  t.h[0] = d.height;
  pool.insert(pool.end(), d.texels, d.texels + size_t(d.width) * d.height * C);
  for (uint32_t l = 1; l < levels; l++) {
    const uint32_t W = t.w[l - 1], H = t.h[l - 1];
    const uint32_t w = std::max(1u, W / 2), h = std::max(1u, H / 2);
    t.off[l] = uint32_t(pool.size());
    t.w[l] = w;
    t.h[l] = h;
    pool.resize(pool.size() + size_t(w) * h * C);
    const float* src = pool.data() + t.off[l - 1];
    float* dst = pool.data() + t.off[l];
    auto wrapi = [&](uint32_t v, uint32_t n) { return d.wrap == RAD_WRAP_REPEAT ? v % n : std::min(v, n - 1); };
    auto tap = [&](uint32_t u, uint32_t v, uint32_t c) {
      const uint32_t pu = wrapi(u, W), pv = wrapi(v, H);
      const uint32_t au = std::min(pu + 1, W - 1), av = std::min(pv + 1, H - 1);
      const float a = src[(size_t(pv) * W + pu) * C + c], b = src[(size_t(pv) * W + au) * C + c];
      const float e = src[(size_t(av) * W + pu) * C + c], f = src[(size_t(av) * W + au) * C + c];
      return (a * 0.5f + b * 0.5f) * 0.5f + (e * 0.5f + f * 0.5f) * 0.5f;
    };
    for (uint32_t j = 0; j < h; j++)
      for (uint32_t i = 0; i < w; i++)
        for (uint32_t c = 0; c < C; c++)
          dst[(size_t(j) * w + i) * C + c] =
              (tap(2 * i, 2 * j, c) + tap(2 * i + 1, 2 * j, c) + tap(2 * i, 2 * j + 1, c) + tap(2 * i + 1, 2 * j + 1, c)) *
              0.25f;
  }
}

void BuildTextureSet(const rad_texture_desc* descs, uint32_t n, std::vector<TexRec>& textures, std::vector<float>& texels) {
  if (n && !descs) throw std::runtime_error("texture_count without textures");
  textures.assign(n, TexRec{});
  texels.clear();
  for (uint32_t i = 0; i < n; i++) {
    const rad_texture_desc& d = descs[i];
    TexRec& t = textures[i];
    t.type = d.type;
    t.channels = d.channels;
    if (d.channels != 1 && d.channels != 3) throw std::runtime_error("texture " + std::to_string(i) + ": channels must be 1 or 3");
    t.tex_a = t.tex_b = -1;
    if (d.type == RAD_TEXTURE_BITMAP) {
      if (!d.texels || d.width == 0 || d.height == 0) throw std::runtime_error("bitmap " + std::to_string(i) + ": empty image");
      if (d.filter < RAD_FILTER_NEAREST || d.filter > RAD_FILTER_ANISOTROPIC) throw std::runtime_error("bitmap: unknown filter");
      t.filter = d.filter;
      t.wrap = d.wrap;
      t.aniso = d.anisotropic_level;
      BuildMipChain(d, t, texels);
    } else if (d.type == RAD_TEXTURE_CHESSBOARD) {
      std::memcpy(t.to_uv, d.to_uv, sizeof(t.to_uv));
      t.remap = d.remap != 0;
      t.color_a = d.channels == 3 ? mk(d.color_a[0], d.color_a[1], d.color_a[2]) : mk(d.color_a[0], 0.f, 0.f);
      t.color_b = d.channels == 3 ? mk(d.color_b[0], d.color_b[1], d.color_b[2]) : mk(d.color_b[0], 0.f, 0.f);
      auto nested = [&](int32_t k) {
        if (k == 0) return -1;
        if (k < 1 || uint32_t(k) > n || descs[k - 1].type != RAD_TEXTURE_BITMAP)
          throw std::runtime_error("chessboard: color_a / color_b must be constants or bitmaps");
        return k - 1;
      };
      t.tex_a = nested(d.tex_a);
      t.tex_b = nested(d.tex_b);
    } else {
      throw std::runtime_error("texture " + std::to_string(i) + ": unknown type");
    }
  }
}

// Texel mean of a texture's level 0 as Texture<T>::Read sees it (a chessboard
// reads 0 as a single texel, chessboard.cpp:29-30, 57-59).
static double TextureMean(const HostScene& h, uint32_t id) {
  const TexRec& t = h.textures[id];
  if (t.type == TEX_CHESSBOARD) return 0.0;
  const float* p = h.texels.data() + t.off[0];
  double sum = 0;
  const size_t n = size_t(t.w[0]) * t.h[0];
  for (size_t i = 0; i < n; i++) sum += t.channels == 3 ? Lum(p + 3 * i) : p[i];
  return sum / double(n);
}

// Textured slots of a BSDF record; plastics take their specular sampling
// weight from the texel means (plastic.cpp:29-42, rough_plastic.cpp:42-55).
static void AttachTextures(const rad_bsdf_desc& d, const HostScene& h, BsdfRec& b) {
  b.tex_mask = 0;
  for (int k = 0; k < RAD_TEX_SLOTS; k++) {
    if (d.tex[k] == 0) continue;
    if (d.tex[k] < 0 || uint32_t(d.tex[k]) > h.textures.size()) throw std::runtime_error("bsdf: texture index out of range");
    const TexRec& t = h.textures[d.tex[k] - 1];
    const bool colour = k == RAD_TEX_REFLECTANCE || k == RAD_TEX_TRANSMITTANCE || k == RAD_TEX_ETA || k == RAD_TEX_K ||
                        k == RAD_TEX_BASE_COLOR || k == RAD_TEX_DIFFUSE || k == RAD_TEX_SPECULAR;
    if ((t.channels == 3) != colour) throw std::runtime_error("bsdf: texture channel count does not match its parameter");
    b.tex_mask |= 1u << k;
    b.tex[k] = uint16_t(d.tex[k] - 1);
  }
  if ((d.type == RAD_BSDF_PLASTIC || d.type == RAD_BSDF_ROUGH_PLASTIC) && (d.tex[RAD_TEX_DIFFUSE] || d.tex[RAD_TEX_SPECULAR])) {
    double dm = d.tex[RAD_TEX_DIFFUSE] ? TextureMean(h, uint32_t(d.tex[RAD_TEX_DIFFUSE] - 1)) : double(Lum(d.diffuse));
    double sm = d.tex[RAD_TEX_SPECULAR] ? TextureMean(h, uint32_t(d.tex[RAD_TEX_SPECULAR] - 1))
                                        : double(Lum(d.specular_reflectance));
    b.spec_weight = float(sm / (dm + sm));
  }
}

// Lookups at a zero footprint (every filter then reads level 0: trilinear and
// anisotropic fall back to the bilinear lookup, bitmap.cpp:82-134), used to
// build the emitters' luminance distributions.
#pragma float_control(push)
#pragma clang fp contract(off)
static f3 HostTexel(const HostScene& h, const TexRec& t, uint32_t x, uint32_t y) {
  const float* p = h.texels.data() + t.off[0] + (size_t(y) * t.w[0] + x) * uint32_t(t.channels);
  return t.channels == 3 ? mk(p[0], p[1], p[2]) : mk(p[0], 0.f, 0.f);
}
static float HostWrap(float v, int32_t wrap) {
  return wrap == RAD_WRAP_REPEAT ? std::min(std::max(v - std::floor(v), 0.f), 1.f) : std::min(std::max(v, 0.f), 1.f);
}
static f3 HostTexLevel0(const HostScene& h, const TexRec& t, float u, float v) {
  const uint32_t W = t.w[0], H = t.h[0];
  const float fu = HostWrap(u, t.wrap) * float(W - 1), fv = HostWrap(v, t.wrap) * float(H - 1);
  const uint32_t pu = uint32_t(fu), pv = uint32_t(fv);
  if (t.filter == RAD_FILTER_NEAREST) return HostTexel(h, t, pu, pv);
  const uint32_t apu = std::min(pu + (fu > float(pu) + 0.5f ? 1u : 0xffffffffu), W - 1);
  const uint32_t apv = std::min(pv + (fv > float(pv) + 0.5f ? 1u : 0xffffffffu), H - 1);
  const float du = std::min(std::fabs(fu - float(pu) - 0.5f), 1.f), dv = std::min(std::fabs(fv - float(pv) - 0.5f), 1.f);
  const f3 a = HostTexel(h, t, pu, pv), b = HostTexel(h, t, apu, pv), c = HostTexel(h, t, pu, apv), d = HostTexel(h, t, apu, apv);
  return (a * (1.f - du) + b * du) * (1.f - dv) + (c * (1.f - du) + d * du) * dv;
}

// ContinuousDistribution2D (distribution.cpp:55-113) over a w x h luminance
// map, stored for the device at light_dist + dist_off.
static void BuildLightDist(const std::vector<float>& lum, uint32_t w, uint32_t h, std::vector<float>& pool, LightRec& r) {
  r.dist_off = uint32_t(pool.size());
  r.dist_w = w;
  r.dist_h = h;
  pool.insert(pool.end(), lum.begin(), lum.end());
  const size_t cdfOff = pool.size(), normOff = cdfOff + size_t(h) * (w + 1), sumOff = normOff + h, mOff = sumOff + h;
  pool.resize(mOff + h + 1);
  for (uint32_t v = 0; v < h; v++) {
    float* cdf = pool.data() + cdfOff + size_t(v) * (w + 1);
    cdf[0] = 0.f;
    for (uint32_t i = 1; i <= w; i++) cdf[i] = cdf[i - 1] + lum[size_t(v) * w + i - 1] / float(w);
    const float sum = cdf[w];
    for (uint32_t i = 1; i <= w; i++) cdf[i] /= sum;
    pool[normOff + v] = float(1.0 / double(sum));
    pool[sumOff + v] = sum;
  }
  float* m = pool.data() + mOff;
  m[0] = 0.f;
  for (uint32_t v = 1; v <= h; v++) m[v] = m[v - 1] + pool[sumOff + v - 1] / float(h);
  r.msum = m[h];
  for (uint32_t v = 1; v <= h; v++) m[v] /= r.msum;
  r.mnorm = float(1.0 / double(r.msum));
}
#pragma float_control(pop)

// Wrappers: the nested index must name another record, and at most two
// wrappers may enclose a base BSDF (the device resolves two levels).
static void ResolveWrappers(std::vector<BsdfRec>& bsdfs) {
  const int32_t n = int32_t(bsdfs.size());
  for (int32_t i = 0; i < n; i++) {
    int32_t j = i, depth = 0;
    while (bsdfs[j].type == BSDF_TWO_SIDE || bsdfs[j].type == BSDF_MASK) {
      int32_t k = bsdfs[j].nested;
      if (k < 0 || k >= n || k == j) throw std::runtime_error("bsdf " + std::to_string(i) + ": nested bsdf out of range");
      if (++depth > 2) throw std::runtime_error("bsdf " + std::to_string(i) + ": more than two nested two_side / mask");
      j = k;
    }
    bsdfs[i].lobes = bsdfs[j].lobes;
  }
  for (int32_t i = 0; i < n; i++) {  // textured: the record or one it wraps reads a texture
    uint32_t any = 0;
    for (int32_t j = i, guard = 0; guard < 3; guard++) {
      any |= bsdfs[j].tex_mask;
      if (bsdfs[j].type != BSDF_TWO_SIDE && bsdfs[j].type != BSDF_MASK) break;
      j = bsdfs[j].nested;
    }
    bsdfs[i].tex_any = any != 0;
  }
}

void AssembleScene(const rad_scene_desc& desc, HostScene& out) {
  out = HostScene();
  BuildTextureSet(desc.textures, desc.texture_count, out.textures, out.texels);
  for (uint32_t i = 0; i < desc.bsdf_count; i++) {
    out.bsdfs.push_back(MakeBsdfRec(desc.bsdfs[i]));
    AttachTextures(desc.bsdfs[i], out, out.bsdfs.back());
  }
  ResolveWrappers(out.bsdfs);
  if (desc.light_count == 0) throw std::runtime_error("scene has no light (Scene::SampleLight needs one)");
  // media
  for (uint32_t i = 0; i < desc.medium_count; i++) {
    const rad_medium_desc& m = desc.media[i];
    MediumRec r{};
    r.type = m.type;
    r.phase = m.phase;
    r.g = m.g;
    if (m.phase == RAD_PHASE_HG && !(m.g > -1.f && m.g < 1.f)) throw std::runtime_error("param g must lie in (-1, 1)");
    r.albedo = mk(m.albedo[0], m.albedo[1], m.albedo[2]);
    r.scale = m.scale;
    if (m.type == RAD_MEDIUM_HOMOGENEOUS) {
      r.sigma_t = mk(m.sigma_t[0] * m.scale, m.sigma_t[1] * m.scale, m.sigma_t[2] * m.scale);
    } else {
      // heterogeneous.cpp:15-29: data lives in [-1,1]^3 of the entity; sampled in [0,1]^3
      M4 A = M4::Identity();
      A.m[0] = A.m[5] = A.m[10] = 2.f;
      A.m[3] = A.m[7] = A.m[11] = -1.f;
      M4 toWorld = M4::From(m.to_world) * A;
      M4 toLocal = toWorld.Inverse();
      r.to_local = toLocal.ToAffine();
      float bb[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
      for (int c = 0; c < 8; c++) GrowBounds(bb, toWorld.Point(mk(c & 1 ? 1.f : 0.f, c & 2 ? 1.f : 0.f, c & 4 ? 1.f : 0.f)));
      r.bbox_lo = mk(bb[0], bb[1], bb[2]);
      r.bbox_hi = mk(bb[3], bb[4], bb[5]);
      float maxValue = -kInf;
      if (m.grid && m.grid_size[0] > 0) {
        if (m.grid_channels != 1 && m.grid_channels != 3) throw std::runtime_error("volume grid needs 1 or 3 channels");
        r.has_grid = 1;
        r.sx = m.grid_size[0];
        r.sy = m.grid_size[1];
        r.sz = m.grid_size[2];
        r.channels = m.grid_channels;
        r.wrap = m.grid_wrap;
        size_t n = size_t(r.sx) * r.sy * r.sz * r.channels;
        r.grid_offset = out.grid.size();
        out.grid.insert(out.grid.end(), m.grid, m.grid + n);
        for (size_t k = 0; k < n; k++) maxValue = std::max(maxValue, m.grid[k]);
      } else {
        r.sigma_t = mk(m.sigma_t[0], m.sigma_t[1], m.sigma_t[2]);
        maxValue = std::max(m.sigma_t[0], std::max(m.sigma_t[1], m.sigma_t[2]));
      }
      r.max_density = maxValue * m.scale;
    }
    out.media.push_back(r);
  }
  out.global_medium = desc.global_medium;
  // lights
  for (uint32_t i = 0; i < desc.light_count; i++) {
    const rad_light_desc& l = desc.lights[i];
    LightRec r{};
    r.type = l.type;
    r.shape = -1;
    r.radiance = mk(l.radiance[0], l.radiance[1], l.radiance[2]);
    r.to_world = M4::From(l.to_world).ToAffine();
    r.tex = -1;
    uint32_t texW = 1, texH = 1;  // Texture::Width / Height (texture.h:30-31); a chessboard is 1 x 1
    if (l.tex != 0) {  // light/area.cpp:20-41, light/skybox.cpp:24-49, light/projection.cpp:21-26
      if (l.type == RAD_LIGHT_POINT) throw std::runtime_error("a point light's intensity cannot be textured");
      if (l.tex < 0 || uint32_t(l.tex) > out.textures.size()) throw std::runtime_error("light: texture index out of range");
      const TexRec& t = out.textures[l.tex - 1];
      if (t.channels != 3) throw std::runtime_error("light: radiance texture must have 3 channels");
      r.tex = l.tex - 1;
      const uint32_t w = t.type == TEX_CHESSBOARD ? 1 : t.w[0], hgt = t.type == TEX_CHESSBOARD ? 1 : t.h[0];
      texW = w;
      texH = hgt;
      r.tex_uniform = w == 1 && hgt == 1;
      if (!r.tex_uniform && l.type != RAD_LIGHT_PROJECTION) {
        std::vector<float> lum(size_t(w) * hgt);
        for (uint32_t y = 0; y < hgt; y++) {
          const float sinTheta = l.type == RAD_LIGHT_SKYBOX ? std::sin((float(y) + 0.5f) / float(hgt) * kPi) : 1.f;
          const float v = float(y) / float(hgt);
          for (uint32_t x = 0; x < w; x++) {
            const f3 c = HostTexLevel0(out, t, float(x) / float(w), v);
            const float lm = (c.x * 0.212671f) + (c.y * 0.715160f) + (c.z * 0.072169f);
            lum[size_t(y) * w + x] = l.type == RAD_LIGHT_SKYBOX ? lm * sinTheta : lm;
          }
        }
        BuildLightDist(lum, w, hgt, out.light_dist, r);
      }
    }
    if (l.type == RAD_LIGHT_SKYBOX) r.to_local = M4::From(l.to_world).Inverse().ToAffine();
    if (l.type == RAD_LIGHT_PROJECTION) {  // projection.cpp:18-45; the aspect is the irradiance image's
      M4 tw = M4::From(l.to_world);
      r.to_local = tw.Inverse().ToAffine();
      const float aspect = float(texW) / float(texH), farC = 10000.f, nearC = 0.0001f;
      const float recip = 1.f / (farC - nearC);
      const float cot = 1.f / std::tan(l.fov * 0.5f * (3.14159265358979323846f / 180.f));
      const float P[16] = {cot, 0, 0, 0, 0, cot, 0, 0, 0, 0, farC * recip, -nearC * farC * recip, 0, 0, 1, 0};
      const float S[16] = {-0.5f, 0, 0, 0, 0, -0.5f * aspect, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
      const float T[16] = {1, 0, 0, -1, 0, 1, 0, -1 / aspect, 0, 0, 1, 0, 0, 0, 0, 1};
      M4 clip = M4::From(S) * M4::From(T) * M4::From(P);
      for (int k = 0; k < 16; k++) r.clip[k] = clip.m[k];
      float ax = tw.m[2], ay = tw.m[6], az = tw.m[10];
      float len = std::sqrt(ax * ax + ay * ay + az * az);
      r.axis = mk(ax / len, ay / len, az / len);
    } else if (l.type != RAD_LIGHT_AREA && l.type != RAD_LIGHT_SKYBOX && l.type != RAD_LIGHT_POINT) {
      throw std::runtime_error("unsupported light type " + std::to_string(l.type));
    }
    out.lights.push_back(r);
    if (l.type == RAD_LIGHT_SKYBOX) out.env_light = int32_t(i);  // the last env light wins (scene.cpp:25-33)
  }
  float wb[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
  // shapes, in geomID order
  for (uint32_t si = 0; si < desc.shape_count; si++) {
    const rad_shape_desc& d = desc.shapes[si];
    ShapeRec s{};
    s.bsdf = d.bsdf;
    s.light = d.light;
    s.med_in = d.medium_inside;
    s.med_out = d.medium_outside;
    if (d.bsdf >= int32_t(desc.bsdf_count) || d.light >= int32_t(desc.light_count) ||
        d.medium_inside >= int32_t(desc.medium_count) || d.medium_outside >= int32_t(desc.medium_count))
      throw std::runtime_error("shape " + std::to_string(si) + " references an out-of-range record");
    if (d.light >= 0) {
      if (desc.lights[d.light].type != RAD_LIGHT_AREA) throw std::runtime_error("only area lights attach to shapes");
      if (out.lights[d.light].tex >= 0 && d.type != RAD_SHAPE_RECTANGLE)
        throw std::runtime_error("textured area lights need a rectangle (the reference's parametric surface, rectangle.cpp:236-250)");
      out.lights[d.light].shape = int32_t(si);
    }
    if (d.type == RAD_SHAPE_CUBE && d.bsdf >= 0 && out.bsdfs[d.bsdf].tex_any)
      throw std::runtime_error("shape " + std::to_string(si) +
                               ": textured BSDF on a cube (the reference never initialises cube texture coordinates, "
                               "cube.cpp:115, 136)");
    M4 tw = M4::From(d.to_world);
    if (d.type == RAD_SHAPE_MESH || d.type == RAD_SHAPE_CUBE) {
      s.type = SHAPE_MESH;
      bool ident = tw.IsIdentity();
      M4 tl = tw.Inverse();
      s.vtx_offset = uint32_t(out.pos.size() / 3);
      s.tri_offset = uint32_t(out.idx.size() / 3);
      uint32_t nv, nt;
      if (d.type == RAD_SHAPE_MESH) {
        if (!d.positions || !d.indices) throw std::runtime_error("mesh without positions / indices");
        nv = d.vertex_count;
        nt = d.triangle_count;
        for (uint32_t v = 0; v < nv; v++) {
          f3 p = mk(d.positions[3 * v], d.positions[3 * v + 1], d.positions[3 * v + 2]);
          if (!ident) p = tw.Point(p);  // mesh.cpp:28-35
          out.pos.insert(out.pos.end(), {p.x, p.y, p.z});
        }
        if (d.normals) {
          s.flags |= SHAPE_HAS_NORMALS;
          for (uint32_t v = 0; v < nv; v++) {
            f3 n = mk(d.normals[3 * v], d.normals[3 * v + 1], d.normals[3 * v + 2]);
            if (!ident) n = tl.TVec(n);  // ApplyNormalToWorld, not renormalised (mesh.cpp:36-42)
            out.nrm.insert(out.nrm.end(), {n.x, n.y, n.z});
          }
        }
        if (d.uvs) {
          s.flags |= SHAPE_HAS_UVS;
          out.uv.insert(out.uv.end(), d.uvs, d.uvs + 2 * size_t(nv));
        }
        for (uint32_t t = 0; t < 3 * nt; t++) {
          if (d.indices[t] >= nv) throw std::runtime_error("mesh index out of range");
          out.idx.push_back(d.indices[t]);
        }
      } else {  // cube.cpp:15-123 (its UVs are never written there; zero here)
        nv = 24;
        nt = 12;
        s.flags |= SHAPE_HAS_NORMALS | SHAPE_HAS_UVS;
        for (int v = 0; v < 24; v++) {
          f3 p = tw.Point(mk(kCubeV[v][0], kCubeV[v][1], kCubeV[v][2]));
          out.pos.insert(out.pos.end(), {p.x, p.y, p.z});
          const float* n = kCubeN[v / 4];
          f3 nn = tl.TVec(mk(n[0], n[1], n[2]));
          out.nrm.insert(out.nrm.end(), {nn.x, nn.y, nn.z});
          out.uv.insert(out.uv.end(), {0.f, 0.f});
        }
        for (int t = 0; t < 12; t++)
          for (int k = 0; k < 3; k++) out.idx.push_back(kCubeT[t][k]);
      }
      // keep position / normal / uv arrays index-aligned with pos for every mesh
      if (!(s.flags & SHAPE_HAS_NORMALS)) out.nrm.resize(out.pos.size(), 0.f);
      if (!(s.flags & SHAPE_HAS_UVS)) out.uv.resize(out.pos.size() / 3 * 2, 0.f);
      s.tri_count = nt;
      // triangle-area CDF accumulated in double (distribution.cpp:19-30, mesh_base.cpp:120-131)
      s.cdf_offset = uint32_t(out.light_cdf.size() / 2);  // (cumulative, area) pairs
      double sum = 0;
      const float* P = out.pos.data() + 3 * size_t(s.vtx_offset);
      const uint32_t* I = out.idx.data() + 3 * size_t(s.tri_offset);
      for (uint32_t t = 0; t < nt; t++) {
        f3 p0 = mk(P[3 * I[3 * t]], P[3 * I[3 * t] + 1], P[3 * I[3 * t] + 2]);
        f3 p1 = mk(P[3 * I[3 * t + 1]], P[3 * I[3 * t + 1] + 1], P[3 * I[3 * t + 1] + 2]);
        f3 p2 = mk(P[3 * I[3 * t + 2]], P[3 * I[3 * t + 2] + 1], P[3 * I[3 * t + 2] + 2]);
        float area = norm(cross(p1 - p0, p2 - p0)) * 0.5f;
        sum += double(area);
        if (d.light >= 0) {
          out.light_cdf.push_back(float(sum));
          out.light_cdf.push_back(area);
        }
        float b[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
        GrowBounds(b, p0);
        GrowBounds(b, p1);
        GrowBounds(b, p2);
        out.prim_bounds.insert(out.prim_bounds.end(), b, b + 6);
        out.prim_records.push_back(F4(p0, AsF(si)));
        out.prim_records.push_back(F4(p1, AsF(t)));
        out.prim_records.push_back(F4(p2, AsF(PRIM_TRI)));
        GrowBounds(wb, p0);
        GrowBounds(wb, p1);
        GrowBounds(wb, p2);
      }
      s.cdf_sum = float(sum);
      s.cdf_norm = float(1.0 / sum);
      s.area = float(sum);
    } else if (d.type == RAD_SHAPE_SPHERE) {  // sphere.cpp:173-190
      s.type = SHAPE_SPHERE;
      float la2 = sqnorm(tw.Vec(mk(1, 0, 0))), lb2 = sqnorm(tw.Vec(mk(0, 1, 0))), lc2 = sqnorm(tw.Vec(mk(0, 0, 1)));
      if (std::fabs(la2 - lb2) > 0.0001f || std::fabs(la2 - lc2) > 0.0001f)
        throw std::runtime_error("Sphere does not support non uniform transformations");
      s.center = tw.Point(mk(d.center[0], d.center[1], d.center[2]));
      s.radius = norm(tw.Vec(mk(0, 0, d.radius)));
      M4 ts = M4::Identity();
      ts.m[0] = ts.m[5] = ts.m[10] = s.radius;
      ts.m[3] = s.center.x;
      ts.m[7] = s.center.y;
      ts.m[11] = s.center.z;
      M4 sw = ts * tw;
      s.to_world = sw.ToAffine();
      s.to_local = sw.Inverse().ToAffine();
      s.area = 4.0f * kPi * sqr(s.radius);
      float b[6];
      b[0] = s.center.x - s.radius;
      b[1] = s.center.y - s.radius;
      b[2] = s.center.z - s.radius;
      b[3] = s.center.x + s.radius;
      b[4] = s.center.y + s.radius;
      b[5] = s.center.z + s.radius;
      out.prim_bounds.insert(out.prim_bounds.end(), b, b + 6);
      out.prim_records.push_back(F4(s.center, s.radius));
      out.prim_records.push_back(make_float4(0, 0, 0, AsF(si)));
      out.prim_records.push_back(make_float4(0, 0, 0, AsF(PRIM_SPHERE)));
      GrowBounds(wb, mk(b[0], b[1], b[2]));
      GrowBounds(wb, mk(b[3], b[4], b[5]));
    } else if (d.type == RAD_SHAPE_RECTANGLE) {  // rectangle.cpp:164-179
      s.type = SHAPE_RECTANGLE;
      M4 t = tw;
      if (d.flip_normal) {
        M4 f = M4::Identity();
        f.m[10] = -1.f;
        t = tw * f;
      }
      M4 tl = t.Inverse();
      s.to_world = t.ToAffine();
      s.to_local = tl.ToAffine();
      f3 dpdu = t.Vec(mk(2, 0, 0)), dpdv = t.Vec(mk(0, 2, 0));
      s.frame_s = dpdu;
      s.frame_t = dpdv;
      s.frame_n = normalize(tl.TVec(mk(0, 0, 1)));
      s.area = norm(cross(dpdu, dpdv));
      float b[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
      GrowBounds(b, t.Point(mk(-1, -1, 0)));
      GrowBounds(b, t.Point(mk(1, -1, 0)));
      GrowBounds(b, t.Point(mk(1, 1, 0)));
      GrowBounds(b, t.Point(mk(-1, 1, 0)));
      out.prim_bounds.insert(out.prim_bounds.end(), b, b + 6);
      out.prim_records.push_back(make_float4(0, 0, 0, 0));
      out.prim_records.push_back(make_float4(0, 0, 0, AsF(si)));
      out.prim_records.push_back(make_float4(0, 0, 0, AsF(PRIM_RECT)));
      GrowBounds(wb, mk(b[0], b[1], b[2]));
      GrowBounds(wb, mk(b[3], b[4], b[5]));
    } else {
      throw std::runtime_error("unknown shape type " + std::to_string(d.type));
    }
    out.shapes.push_back(s);
  }
  for (auto& l : out.lights)
    if (l.type == LIGHT_AREA && l.shape < 0) throw std::runtime_error("area light without a shape");
  for (int a = 0; a < 3; a++) {
    out.world_lo[a] = wb[a];
    out.world_hi[a] = wb[3 + a];
  }
  out.camera_desc = desc.camera;
  out.cam = MakeCameraRec(desc.camera);
}

// Perspective camera record (perspective.cpp:15-70): the inverse of
// scale * translate * perspective with the horizontal field of view.
CameraRec MakeCameraRec(const rad_camera_desc& c) {
  if (c.width <= 0 || c.height <= 0) throw std::runtime_error("camera resolution must be positive");
  CameraRec cam{};
  cam.near_clip = c.near_clip;
  cam.far_clip = c.far_clip;
  cam.width = c.width;
  cam.height = c.height;
  cam.rcp_w = 1.0f / c.width;
  cam.rcp_h = 1.0f / c.height;
  float aspect = c.width / float(c.height);
  float recip = 1.0f / (c.far_clip - c.near_clip);
  float cot = 1.0f / std::tan((c.fov / 2.0f) * (kPi / 180.0f));
  float P[16] = {cot, 0, 0, 0, 0, cot, 0, 0, 0, 0, c.far_clip * recip, -c.near_clip * c.far_clip * recip, 0, 0, 1, 0};
  float S[16] = {-0.5f, 0, 0, 0, 0, -0.5f * aspect, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float T[16] = {1, 0, 0, -1, 0, 1, 0, -1.0f / aspect, 0, 0, 1, 0, 0, 0, 0, 1};
  M4 toClip = M4::From(S) * M4::From(T) * M4::From(P);
  M4 inv = toClip.Inverse();
  std::memcpy(cam.clip_to_camera.m, inv.m, sizeof(inv.m));
  cam.camera_to_world = M4::From(c.to_world).ToAffine();
  // perspective.cpp:55-58: one pixel step on the near plane, in camera space
  const f3 c0 = cam.clip_to_camera.point_h(mk(0.f, 0.f, 0.f));
  cam.dx = cam.clip_to_camera.point_h(mk(1.0f / float(c.width), 0.f, 0.f)) - c0;
  cam.dy = cam.clip_to_camera.point_h(mk(0.f, 1.0f / float(c.height), 0.f)) - c0;
  return cam;
}

}  // namespace rad

uint64_t rad_scene::device_bytes() const {
  return d_nodes.bytes + d_qnodes.bytes + d_prims.bytes + d_shapes.bytes + d_pos.bytes + d_nrm.bytes + d_uv.bytes + d_idx.bytes +
         d_tris.bytes +
         d_bsdfs.bytes + d_lights.bytes + d_cdf.bytes + d_media.bytes + d_grid.bytes + d_tex.bytes + d_texels.bytes +
         d_ldist.bytes + (spare_arena ? spare_arena->bytes : 0);
}
