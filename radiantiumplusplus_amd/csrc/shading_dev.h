// Per-vertex device helpers of the integrators: counter-based RNG, camera ray
// generation, surface-interaction reconstruction, light sampling / pdfs.
#pragma once

#include "bsdf_dev.h"
#include "texture_eval.h"
#include "traverse.h"

namespace rad {

// ---------------------------------------------------------------- RNG
// Independent uniform samples keyed by (seed, request, pixel, sample, dim):
// the stream of a sample never depends on which other work shares the GPU.
RAD_HD uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
RAD_HD uint32_t sample_key(uint32_t seed, uint32_t request, uint32_t pixel, uint32_t sample) {
  uint32_t k = mix32(seed ^ mix32(request * 0x9e3779b9u + 0x85ebca6bu));
  k = mix32(k ^ pixel);
  k = mix32(k ^ (sample * 0xc2b2ae35u + 0x27d4eb2fu));
  return k;
}
RAD_HD float rng_u(uint32_t key, uint32_t dim) {
  uint32_t h = mix32(key ^ mix32(dim * 0x9e3779b9u + 0x632be5abu));
  return float(h >> 8) * 0x1p-24f;  // [0, 1)
}

// ---------------------------------------------------------------- camera (perspective.cpp:73-82)
__device__ inline void camera_ray(const CameraRec& c, float sx, float sy, f3& o, f3& d, float& tmin, float& tmax) {
  f3 ndc = mk(sx * c.rcp_w, sy * c.rcp_h, 0.f);
  f3 nearP = c.clip_to_camera.point_h(ndc);
  f3 dir = normalize(nearP);
  float invZ = 1.0f / dir.z;
  o = c.camera_to_world.point(mk(0.f, 0.f, 0.f));
  d = normalize(c.camera_to_world.vec(dir));
  tmin = c.near_clip * invZ;
  tmax = c.far_clip * invZ;
}

// ---------------------------------------------------------------- surface interaction
struct SurfaceHit {
  f3 p;      // position
  f3 n;      // geometric normal
  Frame sh;  // shading frame
  f3 wi;     // -ray direction in the shading frame
  int32_t shape;
};

RAD_HD f3 load3(const float* a, size_t i) { return mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

// HitShapeRecord::ComputeSurfaceInteraction (interaction.cpp:121-127) with the
// per-shape ComputeInteraction and ComputeShadingFrame (interaction.cpp:45-53).
__device__ inline SurfaceHit surface_hit(const DevScene& s, f3 ro, f3 rd, float t, float u, float v, uint32_t prim,
                                         uint32_t geom) {
  const ShapeRec& sr = s.shapes[geom];
  SurfaceHit h;
  h.shape = int32_t(geom);
  f3 dpdu;
  if (sr.type == SHAPE_MESH) {  // mesh_base.cpp:38-82
    const float4* T = s.tris + 6 * size_t(sr.tri_offset + prim);  // DevScene::tris
    const float4 t0 = T[0], t1 = T[1], t2 = T[2];
    f3 p0 = mk(t0.x, t0.y, t0.z), p1 = mk(t0.w, t1.x, t1.y), p2 = mk(t1.z, t1.w, t2.x);
    float b0 = 1.f - u - v;
    f3 dp0 = p1 - p0, dp1 = p2 - p0;
    h.p = p0 * b0 + (p1 * u + (p2 * v));
    h.n = normalize(cross(dp0, dp1));
    if (sr.flags & SHAPE_HAS_UVS) {
      const float4 t4 = T[4], t5 = T[5];
      float u0 = t4.z, v0 = t4.w;
      float u1 = t5.x, v1 = t5.y;
      float u2 = t5.z, v2 = t5.w;
      float du0 = u1 - u0, dv0 = v1 - v0, du1 = u2 - u0, dv1 = v2 - v0;
      float det = du0 * dv1 - (dv0 * du1);
      float invDet = det == 0.f ? 0.f : 1.f / det;
      dpdu = invDet == 0.f ? splat(0.f) : (dv1 * dp0 - (dv0 * dp1)) * invDet;
    } else {
      f3 y;
      coordinate_system(h.n, dpdu, y);
    }
    if (sr.flags & SHAPE_HAS_NORMALS) {
      const float4 t3 = T[3], t4 = T[4];
      f3 n0 = mk(t2.y, t2.z, t2.w), n1 = mk(t3.x, t3.y, t3.z), n2 = mk(t3.w, t4.x, t4.y);
      f3 sn = n0 * b0 + (n1 * u + (n2 * v));
      h.sh.n = sn * (1.f / sqrtf(sqnorm(sn)));
    } else {
      h.sh.n = h.n;
    }
  } else if (sr.type == SHAPE_SPHERE) {  // sphere.cpp:97 (normal at the hit) and :215-249
    f3 ng = normalize((rd * t + ro) - sr.center);
    h.sh.n = ng;
    h.p = fmadd(ng, splat(sr.radius), sr.center);
    f3 local = sr.to_local.point(h.p);
    dpdu = sr.to_world.vec(mk(-local.y, local.x, 0.f) * (2.f * kPi));
    h.n = ng;
  } else {  // rectangle.cpp:204-219
    f3 p = rd * t + ro;
    f3 tr = mk(sr.to_world.m[3], sr.to_world.m[7], sr.to_world.m[11]);
    h.p = sr.frame_n * dot(tr - p, sr.frame_n) + p;
    h.n = sr.frame_n;
    h.sh.n = sr.frame_n;
    dpdu = sr.frame_s;
  }
  if (is_zero(dpdu)) {
    f3 y;
    coordinate_system(h.sh.n, h.sh.s, y);
  } else {
    h.sh.s = normalize(h.sh.n * (-dot(h.sh.n, dpdu)) + dpdu);
  }
  h.sh.t = cross(h.sh.n, h.sh.s);
  h.wi = h.sh.to_local(-rd);
  return h;
}

// Interaction::OffsetP (interaction.cpp:10-14)
RAD_HD f3 offset_p(f3 p, f3 n, f3 d) {
  float mag = (1.f + maxc(absv(p))) * kRayEps;
  mag = mulsign(mag, dot(n, d));
  return fmadd(splat(mag), n, p);
}

// ---------------------------------------------------------------- lights (scene.cpp:65-118, light/*.cpp)
struct LightSample {
  f3 p, n, dir;
  float dist, pdf;
  f3 li;
  bool delta;
};

// DiscreteDistribution1D::SampleReuse over a mesh emitter's area CDF
// (distribution.cpp:35-53); cdf holds (cumulative, area) pairs.
__device__ inline uint32_t sample_reuse(const float* cdf, uint32_t n, float sum, float norm, float xi, float& rem) {
  float value = xi * sum;
  uint32_t lo = 0, hi = n;  // lower_bound
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (cdf[2 * mid] < value)
      lo = mid + 1;
    else
      hi = mid;
  }
  uint32_t index = min(lo, n - 1);
  float pmf = cdf[2 * index + 1] * norm;
  float c = index == 0 ? 0.f : cdf[2 * (index - 1)] * norm;
  rem = (xi - c) / pmf;
  return index;
}

__device__ inline LightSample sample_area_shape(const DevScene& s, const ShapeRec& sr, f3 refp, float u1, float u2) {
  LightSample ls;
  ls.delta = false;
  if (sr.type == SHAPE_SPHERE) {  // sphere.cpp:265-320
    f3 dcv = sr.center - refp;
    float dc2 = sqnorm(dcv);
    float radAdj = sr.radius * (1.f - kRayEps);
    if (dc2 > sqr(radAdj)) {
      float invDc = 1.f / sqrtf(dc2);
      float sinMax = sr.radius * invDc;
      float sinMax2 = sqr(sinMax);
      float invSinMax = 1.f / sinMax;
      float cosMax = safe_sqrt(1.f - sinMax2);
      float sin2 = sinMax2 > 0.00068523262f ? 1.f - sqr(fmaf(cosMax - 1.f, u1, 1.f)) : sinMax2 * u1;
      float cosT = safe_sqrt(1.f - sin2);
      float cosA = sin2 * invSinMax + cosT * safe_sqrt(fmaf(-sin2, sqr(invSinMax), 1.f));
      float sinA = safe_sqrt(fmaf(-cosA, cosA, 1.f));
      float ang = u2 * (2.f * kPi);
      Frame fr = frame_from_n(dcv * -invDc);
      f3 d = fr.to_world(mk(cosf(ang) * sinA, sinf(ang) * sinA, cosA));
      ls.p = fmadd(d, splat(sr.radius), sr.center);
      ls.n = d;
      ls.dir = ls.p - refp;
      ls.dist = sqrtf(sqnorm(ls.dir));
      ls.dir = ls.dir / ls.dist;
      ls.pdf = 1.f / (2.f * kPi) / (1.f - cosMax);
    } else {
      float z = fmaf(-2.f, u2, 1.f), r = safe_sqrt(fmaf(-z, z, 1.f)), a = 2.f * kPi * u1;
      f3 d = mk(r * cosf(a), r * sinf(a), z);
      ls.p = fmadd(d, splat(sr.radius), sr.center);
      ls.n = d;
      ls.dir = ls.p - refp;
      float dist2 = sqnorm(ls.dir);
      ls.dist = sqrtf(dist2);
      ls.dir = ls.dir / ls.dist;
      ls.pdf = (1.f / sr.area) * dist2 / fabsf(dot(ls.dir, ls.n));
    }
    ls.delta = sr.radius == 0.f;
    return ls;
  }
  float pdfA;
  if (sr.type == SHAPE_RECTANGLE) {  // rectangle.cpp:221-230
    ls.p = sr.to_world.point(mk(u1 * 2.f - 1.f, u2 * 2.f - 1.f, 0.f));
    ls.n = sr.frame_n;
    pdfA = 1.f / sr.area;
  } else {  // mesh_base.cpp:84-112
    float rem;
    uint32_t tri = sample_reuse(s.light_cdf + 2 * size_t(sr.cdf_offset), sr.tri_count, sr.cdf_sum, sr.cdf_norm, u2, rem);
    const uint32_t* I = s.idx + 3 * size_t(sr.tri_offset + tri);
    size_t f0 = sr.vtx_offset + I[0], f1 = sr.vtx_offset + I[1], f2 = sr.vtx_offset + I[2];
    f3 p0 = load3(s.pos, f0), p1 = load3(s.pos, f1), p2 = load3(s.pos, f2);
    f3 e0 = p1 - p0, e1 = p2 - p0;
    float t = safe_sqrt(1.f - u1);
    float bx = 1.f - t, by = t * rem;
    ls.p = e0 * bx + (e1 * by + p0);
    pdfA = sr.cdf_norm;
    if (sr.flags & SHAPE_HAS_NORMALS) {
      f3 n0 = load3(s.nrm, f0), n1 = load3(s.nrm, f1), n2 = load3(s.nrm, f2);
      ls.n = normalize(n0 * (1.f - bx - by) + (n1 * bx + (n2 * by)));
    } else {
      ls.n = normalize(cross(e0, e1));
    }
  }
  // Shape::SampleDirection (shape.cpp:7-18)
  ls.dir = ls.p - refp;
  float d2 = sqnorm(ls.dir);
  ls.dist = sqrtf(d2);
  ls.dir = ls.dir / ls.dist;
  float dp = fabsf(dot(ls.dir, ls.n));
  float x = d2 / dp;
  ls.pdf = pdfA * (isfinite(x) ? x : 0.f);
  return ls;
}

// A radiance texel lookup without differentials, out of line: inlined into
// every shading kernel the filter code costs registers on the common path.
__device__ __noinline__ f3 emitter_texel(const DevScene& s, int32_t tex, float u, float v) {
  return tex_eval(s, uint32_t(tex), TexCoord{u, v, 0.f, 0.f, 0.f, 0.f});
}

// Scene::SampleLightDirection (scene.cpp:87-96)
// Point and projection lights (light/point.cpp:24-38, projection.cpp:40-60),
// kept out of line: rare, and inlined they raise the register count of every
// shading kernel.  A textured projection reads its irradiance image at the
// projected uv with no differentials (interaction.cpp:32-43); `radiance` then
// holds the scale.
//
// TL (here and below): the scene has textured emitters.  With TL false every
// emitter is a constant and the texture code, whose out-of-line calls would
// raise the shading kernels' register counts, is not compiled in.
template <bool TL = true>
__device__ __noinline__ LightSample sample_delta_light(const DevScene& s, const LightRec& L, f3 refp) {
  LightSample ls;
  ls.delta = true;
  ls.pdf = 1.f;
  const f3 pos = mk(L.to_world.m[3], L.to_world.m[7], L.to_world.m[11]);
  if (L.type == LIGHT_POINT) {
    f3 d = pos - refp;
    ls.p = pos;
    ls.n = splat(0.f);
    ls.dist = sqrtf(dot(d, d));
    const float inv = 1.f / ls.dist;
    ls.dir = d * inv;
    ls.li = L.radiance * sqr(inv);
    return ls;
  }
  const f3 local = L.to_local.point(refp);
  const float* c = L.clip;
  const float cw = c[12] * local.x + c[13] * local.y + c[14] * local.z + c[15];
  const float u = (c[0] * local.x + c[1] * local.y + c[2] * local.z + c[3]) / cw;
  const float v = (c[4] * local.x + c[5] * local.y + c[6] * local.z + c[7]) / cw;
  if (!(u >= 0.f && u <= 1.f && v >= 0.f && v <= 1.f && local.z > 0.f)) {
    ls.pdf = 0.f;
    ls.li = splat(0.f);
    ls.p = refp;
    ls.n = splat(0.f);
    ls.dir = mk(0.f, 0.f, 1.f);
    ls.dist = 0.f;
    return ls;
  }
  f3 d = pos - refp;
  ls.p = pos;
  ls.n = L.axis;
  ls.dist = sqrtf(dot(d, d));
  ls.dir = d * (1.f / ls.dist);
  f3 irr = L.radiance;
  if constexpr (TL)
    if (L.tex >= 0) irr = mul(emitter_texel(s, L.tex, u, v), L.radiance);
  ls.li = irr * kPi / (sqr(local.z) * dot(-ls.dir, ls.n));
  return ls;
}

// ---- textured emitters (light/area.cpp:20-148, light/skybox.cpp:24-234)
template <bool TL = true>
__device__ inline f3 emitter_radiance(const DevScene& s, const LightRec& L, float u, float v) {
  if constexpr (!TL) return L.radiance;
  return L.tex < 0 ? L.radiance : emitter_texel(s, L.tex, u, v);
}
// lat-long coordinates of a direction in the skybox frame (skybox.cpp:53-63)
__device__ inline void sky_uv(f3 v, float& u, float& w) {
  float phi = atan2f(v.x, -v.z);
  phi = phi < 0.f ? phi + 2.f * kPi : phi;
  u = phi / (2.f * kPi);
  w = acosf(fminf(fmaxf(v.y, -1.f), 1.f)) / kPi;
}
template <bool TL = true>
__device__ inline f3 skybox_radiance(const DevScene& s, const LightRec& L, f3 worldDir) {
  if (!TL || L.tex < 0) return L.radiance;
  float u, v;
  sky_uv(L.to_local.vec(worldDir), u, v);
  return emitter_radiance(s, L, u, v);
}

// ContinuousDistribution2D::Sample / Pdf (distribution.cpp:85-128) over an
// emitter's luminance map.  The cell is the one below the first cdf entry
// >= xi; the reference forms that index in unsigned arithmetic, so xi at or
// below cdf[0] wraps to the last cell.
// Also returns the continuous position (index + du) / count.
__device__ inline uint32_t dist_cell(const float* cdf, uint32_t count, float xi, float& cont) {
  uint32_t lo = 0, hi = count + 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cdf[mid] < xi)
      lo = mid + 1;
    else
      hi = mid;
  }
  const uint32_t i = lo == 0 ? count - 1 : min(lo - 1, count - 1);
  float du = xi - cdf[i];
  if ((cdf[i + 1] - cdf[i]) > 0.f) du /= (cdf[i + 1] - cdf[i]);
  cont = (float(i) + du) / float(count);
  return i;
}
struct DistSample {
  uint32_t u, v;
  float pdf;
  float cu, cv;  // continuous position
};
__device__ inline DistSample light_dist_sample(const DevScene& s, const LightRec& L, float xu, float xv) {
  const uint32_t w = L.dist_w, h = L.dist_h;
  const float* lum = s.light_dist + L.dist_off;
  const float* ccdf = lum + size_t(w) * h;
  const float* cnorm = ccdf + size_t(h) * (w + 1);
  const float* rsum = cnorm + h;
  const float* mcdf = rsum + h;
  DistSample d;
  d.v = dist_cell(mcdf, h, xv, d.cv);
  d.u = dist_cell(ccdf + size_t(d.v) * (w + 1), w, xu, d.cu);
  d.pdf = (lum[size_t(d.v) * w + d.u] * cnorm[d.v]) * (rsum[d.v] * L.mnorm);
  return d;
}
__device__ inline float light_dist_pdf(const DevScene& s, const LightRec& L, float u, float v) {
  const uint32_t w = L.dist_w, h = L.dist_h;
  const float fu = u * float(w), fv = v * float(h);
  const uint32_t iu = fu >= float(w) ? w - 1 : (fu > 0.f ? uint32_t(fu) : 0u);
  const uint32_t iv = fv >= float(h) ? h - 1 : (fv > 0.f ? uint32_t(fv) : 0u);
  return s.light_dist[L.dist_off + size_t(iv) * w + iu] / L.msum;
}

// Textured area light (more than one texel) on its rectangle: sampled at the
// corner uv of a luminance-distributed texel (area.cpp:56-73).  Its
// SampleDirection record comes from SurfaceInteraction::ToDsr, whose density
// is 0 (interaction.cpp:59-75): with quirks on the sample is discarded and the
// light is reached by BSDF sampling only.  Quirks off: the continuous
// position inside the texel with the PdfDirection density.
__device__ __noinline__ LightSample sample_area_textured(const DevScene& s, const LightRec& L, f3 refp, float u1,
                                                         float u2, bool quirks) {
  LightSample ls;
  ls.delta = false;
  const DistSample d = light_dist_sample(s, L, u1, u2);
  const ShapeRec& sr = s.shapes[L.shape];
  const float u = quirks ? float(d.u) / float(L.dist_w) : d.cu, v = quirks ? float(d.v) / float(L.dist_h) : d.cv;
  ls.p = sr.to_world.point(mk(u * 2.f - 1.f, v * 2.f - 1.f, 0.f));
  ls.n = sr.frame_n;
  ls.dir = ls.p - refp;
  ls.dist = sqrtf(sqnorm(ls.dir));
  ls.dir = ls.dir / ls.dist;
  ls.li = dot(-ls.dir, ls.n) >= 0.f ? emitter_radiance(s, L, u, v) : splat(0.f);
  if (!(d.pdf > 0.f) || quirks) {
    ls.pdf = 0.f;
    return ls;
  }
  const float pa = light_dist_pdf(s, L, u, v);
  ls.pdf = pa * sqr(ls.dist) / (norm(cross(sr.frame_s, sr.frame_t)) * -dot(ls.dir, ls.n));
  return ls;
}

// Lat-long map of more than one texel (skybox.cpp:92-108, 215-234): the
// texel's corner, or (quirks off) the continuous position inside it.
__device__ __noinline__ LightSample sample_skybox_textured(const DevScene& s, const LightRec& L, f3 refp, float u1,
                                                           float u2, bool quirks) {
  LightSample ls;
  ls.delta = false;
  const float radius = fmaxf(s.world_radius, norm(refp - s.world_center));
  ls.dist = 2.f * radius;
  ls.pdf = 0.f;
  ls.li = splat(0.f);
  ls.p = refp;
  ls.n = splat(0.f);
  ls.dir = mk(0.f, 0.f, 1.f);
  const DistSample d = light_dist_sample(s, L, u1, u2);
  if (!(d.pdf > 0.f)) return ls;
  const float y = quirks ? float(d.v) / float(L.dist_h) : d.cv, x = quirks ? float(d.u) / float(L.dist_w) : d.cu;
  const float theta = y * kPi, phi = x * 2.f * kPi;
  const float sinT = sinf(theta);
  if (sinT == 0.f) return ls;
  const f3 sph = mk(cosf(phi) * sinT, sinf(phi) * sinT, cosf(theta));  // math_ext.h:108-112
  const f3 wd = L.to_world.vec(mk(sph.y, sph.z, -sph.x));
  ls.p = wd * ls.dist + refp;
  ls.n = -wd;
  ls.dir = wd;
  ls.pdf = d.pdf / (2.f * kPi * kPi * sinT);
  ls.li = emitter_radiance(s, L, x, y);
  return ls;
}

template <bool TL = true>
__device__ inline LightSample sample_light(const DevScene& s, f3 refp, float ul, float u1, float u2, bool quirks) {
  uint32_t index = 0;
  float lpdf = 1.f;
  if (s.light_count >= 2) {
    index = min(uint32_t(ul * float(s.light_count)), s.light_count - 1);
    lpdf = s.light_pdf;
  }
  const LightRec& L = s.lights[index];
  LightSample ls;
  if (TL && L.type == LIGHT_SKYBOX && L.tex >= 0 && !L.tex_uniform) {
    if constexpr (TL) ls = sample_skybox_textured(s, L, refp, u1, u2, quirks);
  } else if (L.type == LIGHT_SKYBOX) {  // skybox.cpp:65-88 (one-texel map)
    float radius = fmaxf(s.world_radius, norm(refp - s.world_center));
    float dist = 2.f * radius;
    float z = fmaf(-2.f, u2, 1.f), r = safe_sqrt(fmaf(-z, z, 1.f)), a = 2.f * kPi * u1;
    f3 dir = normalize(mk(r * cosf(a), r * sinf(a), z));
    f3 wd = L.to_world.vec(dir);
    ls.p = wd * dist + refp;
    ls.n = -wd;
    ls.dir = wd;
    ls.dist = dist;
    ls.pdf = 1.f / (4.f * kPi);
    ls.delta = false;
    if (!TL || L.tex < 0) {
      ls.li = L.radiance;
    } else {
      float tu, tv;
      sky_uv(dir, tu, tv);
      ls.li = emitter_radiance<TL>(s, L, tu, tv);
    }
  } else if (L.type == LIGHT_POINT || L.type == LIGHT_PROJECTION) {
    ls = sample_delta_light<TL>(s, L, refp);
  } else if (TL && L.tex >= 0 && !L.tex_uniform) {
    if constexpr (TL) ls = sample_area_textured(s, L, refp, u1, u2, quirks);
  } else {  // light/area.cpp:50-75 (a one-texel texture is read at the rectangle's uv = xi)
    ls = sample_area_shape(s, s.shapes[L.shape], refp, u1, u2);
    ls.li = dot(-ls.dir, ls.n) >= 0.f ? emitter_radiance<TL>(s, L, u1, u2) : splat(0.f);
  }
  ls.pdf *= lpdf;
  return ls;
}

// PdfDirection of a textured emitter (area.cpp:77-91, skybox.cpp:110-124).
__device__ __noinline__ float textured_pdf_dir(const DevScene& s, const LightRec& L, f3 n_shading, f3 dir, float dist,
                                               float hu, float hv, bool quirks) {
  if (L.type == LIGHT_SKYBOX) {  // the affine inverse, translation included (Q15), unless quirks are off
    const f3 v = quirks ? L.to_local.point(dir) : L.to_local.vec(dir);
    const float theta = acosf(fminf(fmaxf(v.y, -1.f), 1.f));
    const float sinT = sinf(theta);
    if (sinT == 0.f) return 0.f;
    float u, w;
    sky_uv(v, u, w);
    return light_dist_pdf(s, L, u, w) / (2.f * kPi * kPi * sinT);
  }
  const float pa = light_dist_pdf(s, L, hu, hv);
  if (pa <= 0.f) return 0.f;
  const ShapeRec& sr = s.shapes[L.shape];
  return pa * sqr(dist) / (norm(cross(sr.frame_s, sr.frame_t)) * -dot(dir, n_shading));
}

// Light::PdfDirection * 1/N for a light hit from `refp` (scene.cpp:98-103);
// (hu, hv): the hit's texture coordinates on a textured rectangle emitter.
template <bool TL = true>
__device__ inline float light_pdf_dir(const DevScene& s, int32_t light, f3 refp, f3 p, f3 n_shading, f3 dir,
                                      float dist, float hu, float hv, bool quirks) {
  const LightRec& L = s.lights[light];
  if constexpr (TL)
    if (L.tex >= 0 && !L.tex_uniform) return textured_pdf_dir(s, L, n_shading, dir, dist, hu, hv, quirks) * s.light_pdf;
  if (L.type == LIGHT_SKYBOX) return (1.f / (4.f * kPi)) * s.light_pdf;
  if (!(dot(-dir, n_shading) >= 0.f)) return 0.f;  // area.cpp:77-80
  const ShapeRec& sr = s.shapes[L.shape];
  float pdf;
  if (sr.type == SHAPE_SPHERE) {  // sphere.cpp:322-328
    float sinA = sr.radius * (1.f / norm(sr.center - refp));
    float cosA = safe_sqrt(1.f - sinA * sinA);
    pdf = sinA < kOneMinusEps ? 1.f / (2.f * kPi) / (1.f - cosA) : (1.f / sr.area) * sqr(dist) / fabsf(dot(dir, n_shading));
  } else {  // shape.cpp:20-25
    float pa = sr.type == SHAPE_RECTANGLE ? 1.f / sr.area : sr.cdf_norm;
    float dp = fabsf(dot(dir, n_shading));
    pdf = pa * (dp == 0.f ? 0.f : sqr(dist) / dp);
  }
  return pdf * s.light_pdf;
}

RAD_HD float mis_weight(float a, float b) {  // power heuristic (path.cpp:231-236)
  a *= a;
  b *= b;
  float w = a / (a + b);
  return isfinite(w) ? w : 0.f;
}

}  // namespace rad
