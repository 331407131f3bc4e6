// Device ray traversal over the 4-wide BVH: closest hit
// (Embree::RayIntersectPreliminary, embree.cpp:87-111) and occlusion
// (Embree::RayIntersect, embree.cpp:79-85).
#pragma once

#include "scene_dev.h"

namespace rad {

struct Hit {
  float t, u, v;
  uint32_t prim;
  uint32_t geom;  // 0xffffffff = miss
};

// Products of the primitive tests with every rounding spelled out (explicit
// fused multiply-adds in a fixed association), so the device and the host
// oracle produce bit-identical t, u, v whatever contraction the compiler
// applies elsewhere.
__device__ __forceinline__ float dotx(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ f3 crossx(f3 a, f3 b) {
  return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}

// Moller-Trumbore in the form Embree 3 applies to its TriangleM primitives:
// e1 = v0 - v1, e2 = v2 - v0, Ng = e2 x e1, C = v0 - O, R = C x D,
// accept den != 0, U, V >= 0, U + V <= |den|, |den|*tnear < T <= |den|*tfar.
__device__ __forceinline__ bool isect_tri(f3 v0, f3 v1, f3 v2, f3 o, f3 d, float tnear, float tfar, float& t,
                                          float& u, float& v) {
  f3 e1 = v0 - v1;
  f3 e2 = v2 - v0;
  f3 ng = crossx(e2, e1);
  f3 c = v0 - o;
  f3 r = crossx(c, d);
  float den = dotx(ng, d);
  float absDen = fabsf(den);
  float U = dotx(r, e2);
  float V = dotx(r, e1);
  if (den < 0.f) {
    U = -U;
    V = -V;
  }
  if (!(den != 0.f && U >= 0.f && V >= 0.f && U + V <= absDen)) return false;
  float T = dotx(ng, c);
  if (den < 0.f) T = -T;
  if (!(absDen * tnear < T && T <= absDen * tfar)) return false;
  t = T / absDen;
  u = U / absDen;
  v = V / absDen;
  return true;
}

// Sphere user geometry (sphere.cpp:40-54 with SolveQuadratic, math_ext.h:79-97).
__device__ __forceinline__ bool isect_sphere(f3 center, float radius, f3 ro, f3 rd, float near, float far, float& t) {
  f3 o = ro - center;
  float a = dotx(rd, rd);
  float b = 2.0f * dotx(o, rd);
  float c = fmaf(-radius, radius, dotx(o, o));
  bool linearCase = a == 0.0f;
  bool validLinear = linearCase && b != 0.0f;
  float x0 = -c / b, x1 = -c / b;
  float discrim = fmaf(b, b, -((4.0f * a) * c));
  bool validQuadratic = !linearCase && (discrim >= 0.0f);
  if (validQuadratic) {
    float rootDiscrim = sqrtf(discrim);
    float temp = -0.5f * (b + copysignf(rootDiscrim, b));
    float x0p = temp / a;
    float x1p = c / temp;
    float x0m = fminf(x0p, x1p);
    float x1m = fmaxf(x0p, x1p);
    x0 = linearCase ? x0 : x0m;
    x1 = linearCase ? x0 : x1m;
  }
  bool found = validLinear || validQuadratic;
  bool outBounds = !(x0 <= far && x1 >= near);
  bool inBounds = x0 < near && x1 > far;
  if (!(found && !outBounds && !inBounds)) return false;
  t = x0 < 0.f ? x1 : x0;
  return true;
}

// Rectangle user geometry (rectangle.cpp:34-46): object-space plane z = 0,
// |x|, |y| <= 1, with an unnormalised object-space direction.
__device__ __forceinline__ bool isect_rect(const Affine& toObj, f3 ro, f3 rd, float near, float far, float& t,
                                           float& u, float& v) {
  const float* m = toObj.m;
  f3 o = mk(fmaf(m[0], ro.x, fmaf(m[1], ro.y, fmaf(m[2], ro.z, m[3]))),
            fmaf(m[4], ro.x, fmaf(m[5], ro.y, fmaf(m[6], ro.z, m[7]))),
            fmaf(m[8], ro.x, fmaf(m[9], ro.y, fmaf(m[10], ro.z, m[11]))));
  f3 d = mk(fmaf(m[0], rd.x, fmaf(m[1], rd.y, m[2] * rd.z)), fmaf(m[4], rd.x, fmaf(m[5], rd.y, m[6] * rd.z)),
            fmaf(m[8], rd.x, fmaf(m[9], rd.y, m[10] * rd.z)));
  float tt = -o.z / d.z;
  f3 local = mk(fmaf(d.x, tt, o.x), fmaf(d.y, tt, o.y), fmaf(d.z, tt, o.z));
  if (!(tt >= near && tt <= far && fabsf(local.x) <= 1.f && fabsf(local.y) <= 1.f)) return false;
  t = tt;
  u = local.x;
  v = local.y;
  return true;
}

__device__ __forceinline__ f3 xyz(float4 a) { return mk(a.x, a.y, a.z); }

// Test one primitive record; on a hit narrows `tfar` and fills `h`.
__device__ __forceinline__ bool isect_prim(const DevScene& s, uint32_t p, f3 o, f3 d, float tnear, float& tfar,
                                           Hit& h) {
  const float4* rec = s.prims + 3 * size_t(p);
  float4 a = rec[0], b = rec[1], c = rec[2];
  uint32_t kind = __float_as_uint(c.w);
  float t, u = 0.f, v = 0.f;
  if (kind == PRIM_TRI) {
    if (!isect_tri(xyz(a), xyz(b), xyz(c), o, d, tnear, tfar, t, u, v)) return false;
    h.geom = __float_as_uint(a.w);
    h.prim = __float_as_uint(b.w);
  } else if (kind == PRIM_SPHERE) {
    if (!isect_sphere(xyz(a), a.w, o, d, tnear, tfar, t)) return false;
    h.geom = __float_as_uint(b.w);
    h.prim = 0;
  } else {
    uint32_t g = __float_as_uint(b.w);
    if (!isect_rect(s.shapes[g].to_local, o, d, tnear, tfar, t, u, v)) return false;
    h.geom = g;
    h.prim = 0;
  }
  h.t = t;
  h.u = u;
  h.v = v;
  tfar = t;
  return true;
}

constexpr int kStackDepth = 64;

// ANY = true: return at the first accepted primitive (occlusion).
template <bool ANY>
__device__ __forceinline__ bool trace(const DevScene& s, f3 o, f3 d, float tmin, float tmax, Hit& h,
                                      uint32_t* visits = nullptr) {
  f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  f3 oinv = mk(o.x * inv.x, o.y * inv.y, o.z * inv.z);
  uint32_t stack[kStackDepth];
  int sp = 0;
  uint32_t node = 0;
  float tfar = tmax;
  bool hit = false;
  uint32_t nvisit = 0;
  const float4* base = reinterpret_cast<const float4*>(s.nodes);
  for (;;) {
    const float4* q = base + 8 * size_t(node);
    float4 lx = q[0], hx = q[1], ly = q[2], hy = q[3], lz = q[4], hz = q[5];
    int4 ch = *reinterpret_cast<const int4*>(q + 6);
    uint4 cn = *reinterpret_cast<const uint4*>(q + 7);
    nvisit++;
    float lo_x[4] = {lx.x, lx.y, lx.z, lx.w}, hi_x[4] = {hx.x, hx.y, hx.z, hx.w};
    float lo_y[4] = {ly.x, ly.y, ly.z, ly.w}, hi_y[4] = {hy.x, hy.y, hy.z, hy.w};
    float lo_z[4] = {lz.x, lz.y, lz.z, lz.w}, hi_z[4] = {hz.x, hz.y, hz.z, hz.w};
    int cc[4] = {ch.x, ch.y, ch.z, ch.w};
    uint32_t cnt[4] = {cn.x, cn.y, cn.z, cn.w};
    float tn[4];
    bool m[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      float t0x = fmaf(lo_x[i], inv.x, -oinv.x), t1x = fmaf(hi_x[i], inv.x, -oinv.x);
      float t0y = fmaf(lo_y[i], inv.y, -oinv.y), t1y = fmaf(hi_y[i], inv.y, -oinv.y);
      float t0z = fmaf(lo_z[i], inv.z, -oinv.z), t1z = fmaf(hi_z[i], inv.z, -oinv.z);
      float tnear = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
      float tfb = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tfar));
      // slack for the rounding of the fused slab arithmetic
      tnear = tnear - fabsf(tnear) * 4.8e-7f;
      tfb = tfb + fabsf(tfb) * 4.8e-7f;
      m[i] = cc[i] >= 0 && tnear <= tfb;
      tn[i] = tnear;
    }
    // leaves first (nearest-first among them does not matter for correctness)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (m[i] && cnt[i] > 0) {
        uint32_t first = uint32_t(cc[i]);
        for (uint32_t k = 0; k < cnt[i]; k++) {
          if (isect_prim(s, first + k, o, d, tmin, tfar, h)) {
            hit = true;
            if (ANY) {
              if (visits) *visits = nvisit;
              return true;
            }
          }
        }
      }
    }
    // inner children, nearest first
    float bd = kInf;
    int bi = -1;
    int ni = 0;
    uint32_t inner[4];
    float dist[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (m[i] && cnt[i] == 0 && tn[i] <= tfar) {
        inner[ni] = uint32_t(cc[i]);
        dist[ni] = tn[i];
        ni++;
      }
    }
    if (ni > 0) {
      // order by distance: push the far ones, descend into the nearest
      for (int a = 1; a < ni; a++) {
        for (int b = a; b > 0 && dist[b] > dist[b - 1]; b--) {
          float tdist = dist[b];
          dist[b] = dist[b - 1];
          dist[b - 1] = tdist;
          uint32_t tn2 = inner[b];
          inner[b] = inner[b - 1];
          inner[b - 1] = tn2;
        }
      }
      // dist is now descending; the last entry is the nearest
      for (int a = 0; a < ni - 1; a++) {
        if (sp < kStackDepth) stack[sp++] = inner[a];
      }
      node = inner[ni - 1];
      (void)bd;
      (void)bi;
      continue;
    }
    if (sp == 0) break;
    node = stack[--sp];
  }
  if (visits) *visits = nvisit;
  return hit;
}

}  // namespace rad
