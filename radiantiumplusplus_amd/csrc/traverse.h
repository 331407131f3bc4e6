// Device ray traversal over the 4-wide BVH (its 64-byte quantised copy):
// closest hit
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
// Traversal keeps only t = T / |den| of the closest hit; tri_uv recomputes
// U, V and |den| with the same operations once the ray is finished.
__device__ __forceinline__ bool isect_tri(f3 v0, f3 v1, f3 v2, f3 o, f3 d, float tnear, float tfar, float& t) {
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
  return true;
}

__device__ __forceinline__ void tri_uv(f3 v0, f3 v1, f3 v2, f3 o, f3 d, float& u, float& v) {
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
  u = U / absDen;
  v = V / absDen;
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
                                           float& u, float& v) {  // u, v: the local hit point
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

// Test one primitive record; on a hit narrows `tfar` to its distance.
__device__ __forceinline__ bool isect_prim(const DevScene& s, uint32_t p, f3 o, f3 d, float tnear, float& tfar) {
  const float4* rec = s.prims + 3 * size_t(p);
  const float4 c = rec[2];
  const uint32_t kind = __float_as_uint(c.w);
  float t;
  if (kind == PRIM_TRI) {
    const float4 a = rec[0], b = rec[1];
    if (!isect_tri(xyz(a), xyz(b), xyz(c), o, d, tnear, tfar, t)) return false;
  } else if (kind == PRIM_SPHERE) {
    const float4 a = rec[0];
    if (!isect_sphere(xyz(a), a.w, o, d, tnear, tfar, t)) return false;
  } else {
    float u, v;
    if (!isect_rect(s.shapes[__float_as_uint(rec[1].w)].to_local, o, d, tnear, tfar, t, u, v)) return false;
  }
  tfar = t;
  return true;
}

// The hit record of primitive p at distance t (the closest hit the
// traversal kept): geometry and primitive ids from the record, u, v
// recomputed with the arithmetic of the test that accepted it.
__device__ __forceinline__ void prim_hit(const DevScene& s, uint32_t p, f3 o, f3 d, float t, Hit& h) {
  const float4* rec = s.prims + 3 * size_t(p);
  const float4 a = rec[0], b = rec[1], c = rec[2];
  const uint32_t kind = __float_as_uint(c.w);
  h.t = t;
  h.u = 0.f;
  h.v = 0.f;
  if (kind == PRIM_TRI) {
    tri_uv(xyz(a), xyz(b), xyz(c), o, d, h.u, h.v);
    h.geom = __float_as_uint(a.w);
    h.prim = __float_as_uint(b.w);
  } else if (kind == PRIM_SPHERE) {
    h.geom = __float_as_uint(b.w);
    h.prim = 0;
  } else {
    const uint32_t g = __float_as_uint(b.w);
    float tt;
    isect_rect(s.shapes[g].to_local, o, d, -kInf, kInf, tt, h.u, h.v);
    h.geom = g;
    h.prim = 0;
  }
}

// ---------------------------------------------------------------- traversal
// Each lane keeps its traversal stack as (node, entry distance) pairs: the
// newest kLdsStack entries in LDS (a ring, lane-interleaved so a wave's
// accesses hit distinct banks), older ones spilled to a private overflow
// array that only deep traversals touch.
constexpr int kStackDepth = int(kTraceStackEntries);
constexpr int kLdsStack = 8;
constexpr int kTraceBlock = 256;  // threads per block of every traversal kernel

// LDS pointers spelled with their address space, so stack accesses compile
// to ds_read / ds_write rather than flat accesses.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) float lds_f32;

// Per-lane private overflow of a traversal stack (entries below the LDS ring).
struct TraceOverflow {
  uint32_t node[kStackDepth - kLdsStack];
  float dist[kStackDepth - kLdsStack];
};

// The struct holds no arrays, so its fields stay in registers; only the
// overflow array, indexed by the stack depth, lives in private memory.
struct TraceStack {
  lds_u32* node;  // LDS, entry k of this lane at node[k * kTraceBlock]
  lds_f32* dist;
  uint32_t* ov_node;  // TraceOverflow of this lane
  float* ov_dist;
  int sp, lo;  // entries [0, lo) live in the overflow array, [lo, sp) in LDS

  __device__ __forceinline__ void push(uint32_t n, float d) {
    if (sp >= kStackDepth) return;  // unreachable: the builders refuse trees that need more entries
    if (sp - lo == kLdsStack) {
      int k = (lo & (kLdsStack - 1)) * kTraceBlock;
      ov_node[lo] = node[k];
      ov_dist[lo] = dist[k];
      lo++;
    }
    int k = (sp & (kLdsStack - 1)) * kTraceBlock;
    node[k] = n;
    dist[k] = d;
    sp++;
  }
  __device__ __forceinline__ void pop(uint32_t& n, float& d) {
    sp--;
    if (sp < lo) {
      lo = sp;
      n = ov_node[sp];
      d = ov_dist[sp];
    } else {
      int k = (sp & (kLdsStack - 1)) * kTraceBlock;
      n = node[k];
      d = dist[k];
    }
  }
};

struct TraceRay {
  f3 o, d;
  f3 inv;           // 1 / d, components below 1e-18 in magnitude raised to it (sign kept)
  bool nx, ny, nz;  // direction components negative: the near plane is the box's high bound
  float tmin, tfar;  // tfar: the closest hit's distance once hit is set
  uint32_t node;
  uint32_t hitp;  // primitive record of the closest hit
  bool hit;
  Hit h;  // filled by trace_finish
};

// Fills the hit record of a finished closest-hit traversal.
__device__ __forceinline__ void trace_finish(const DevScene& s, TraceRay& r) {
  if (r.hit) {
    prim_hit(s, r.hitp, r.o, r.d, r.tfar, r.h);
  } else {
    r.h.t = r.tfar;
    r.h.u = r.h.v = 0.f;
    r.h.prim = 0;
    r.h.geom = 0xffffffffu;
  }
}

// A closest-hit traversal interrupted at a launch's visit budget: the node
// to visit next, the pending stack, the closest hit so far.  Resuming it
// replays nothing, so the result is the one an uninterrupted traversal gives.
struct ResumeState {
  uint32_t node, sp, hit, hitp;
  float tfar;
  uint32_t pad[7];
  uint32_t stack_node[kStackDepth];
  float stack_dist[kStackDepth];
};

__device__ inline void save_trace(ResumeState& rs, const TraceRay& r, const TraceStack& st) {
  rs.node = r.node;
  rs.sp = uint32_t(st.sp);
  rs.hit = r.hit ? 1u : 0u;
  rs.hitp = r.hitp;
  rs.tfar = r.tfar;
  for (int k = 0; k < st.sp; k++) {
    if (k < st.lo) {
      rs.stack_node[k] = st.ov_node[k];
      rs.stack_dist[k] = st.ov_dist[k];
    } else {
      const int j = (k & (kLdsStack - 1)) * kTraceBlock;
      rs.stack_node[k] = st.node[j];
      rs.stack_dist[k] = st.dist[j];
    }
  }
}

// After trace_begin on the same ray.
__device__ inline void restore_trace(const ResumeState& rs, TraceRay& r, TraceStack& st) {
  r.node = rs.node;
  r.hit = rs.hit != 0;
  r.hitp = rs.hitp;
  r.tfar = rs.tfar;
  const int n = min(int(rs.sp), kStackDepth);
  for (int k = 0; k < n; k++) st.push(rs.stack_node[k], rs.stack_dist[k]);
}

__device__ __forceinline__ void trace_begin(TraceRay& r, TraceStack& st, f3 o, f3 d, float tmin, float tmax) {
  r.o = o;
  r.d = d;
  // A zero direction component would make the slab arithmetic produce
  // inf - inf; a tiny one of the same sign keeps every slab finite.
  const auto safe = [](float x) { return fabsf(x) < 1e-18f ? copysignf(1e-18f, x) : x; };
  r.inv = mk(1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z));
  r.nx = r.inv.x < 0.f;
  r.ny = r.inv.y < 0.f;
  r.nz = r.inv.z < 0.f;
  r.tmin = tmin;
  r.tfar = tmax;
  r.node = 0;
  r.hit = false;
  r.hitp = 0xffffffffu;
  st.sp = 0;
  st.lo = 0;
}

__device__ __forceinline__ void cswap(float& da, uint32_t& na, float& db, uint32_t& nb) {
  bool sw = db < da;
  float td = sw ? db : da;
  uint32_t tn = sw ? nb : na;
  db = sw ? da : db;
  nb = sw ? na : nb;
  da = td;
  na = tn;
}

// Per-wave LDS of the pooled primitive tests: a counter, kPoolSize
// (primitive, owner lane) entries and one closest-hit key per lane.
constexpr uint32_t kPoolSize = kTracePool;
struct PrimPool {
  lds_u32* ctr;    // entries posted in this node step
  lds_u32* prim;   // [kPoolSize]
  lds_u32* owner;  // [kPoolSize]
  __attribute__((address_space(3))) unsigned long long* key;  // this wave's 64 lanes
};

// Orders the wave's LDS accesses around the pooled tests: LDS executes one
// wave's instructions in order, so only the compiler needs holding back.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Visit one node: test its four child boxes, push its hit inner children
// (far first) and descend into the nearest, and intersect the primitives of
// its hit leaf children; the closest hit is the minimum of (t, primitive)
// over the accepted primitives.  POOL: the lanes of a wave test their leaf
// primitives as one pool: each lane posts its (primitive, lane) entries to
// LDS, every lane taking part in this step tests entries of the pool against
// the owner's ray, and a 64-bit LDS minimum returns each owner its closest
// hit; every lane whose ray is in flight then calls it together (the pool
// reads other lanes' registers).  Without POOL each lane tests its own
// primitives.  Returns true when the ray is finished.
template <bool ANY, bool POOL>
__device__ __forceinline__ bool trace_step(const DevScene& s, TraceRay& r, TraceStack& st, const PrimPool& pool) {
  // one 64-byte quantised node: origin + steps, 8-bit child bounds, counts, children
  const uint4* q = reinterpret_cast<const uint4*>(s.qnodes) + 4 * size_t(r.node);
  const uint4 hdr = q[0], bxy = q[1], bz = q[2];
  const int4 ch = *reinterpret_cast<const int4*>(q + 3);
  // A child plane at 8-bit offset q along x lies at origin + q * step; its ray
  // parameter is q * (step / d) + (origin - o) / d.  step is a power of two,
  // so the first factor is exact, and the near plane of each axis is chosen
  // once per node by the direction's sign.
  const float ax = quant_step(hdr.w) * r.inv.x, ay = quant_step(hdr.w >> 8) * r.inv.y,
              az = quant_step(hdr.w >> 16) * r.inv.z;
  const float bx = (__uint_as_float(hdr.x) - r.o.x) * r.inv.x, by = (__uint_as_float(hdr.y) - r.o.y) * r.inv.y,
              bz0 = (__uint_as_float(hdr.z) - r.o.z) * r.inv.z;
  const uint32_t nearx = r.nx ? bxy.y : bxy.x, farx = r.nx ? bxy.x : bxy.y;
  const uint32_t neary = r.ny ? bxy.w : bxy.z, fary = r.ny ? bxy.z : bxy.w;
  const uint32_t nearz = r.nz ? bz.y : bz.x, farz = r.nz ? bz.x : bz.y;
  const int cc[4] = {ch.x, ch.y, ch.z, ch.w};
  uint32_t cnt[4];
  float tn[4];
  bool m[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int sh = 8 * i;
    cnt[i] = (bz.z >> sh) & 0xffu;
    const float tnx = fmaf(float((nearx >> sh) & 0xffu), ax, bx), tfx = fmaf(float((farx >> sh) & 0xffu), ax, bx);
    const float tny = fmaf(float((neary >> sh) & 0xffu), ay, by), tfy = fmaf(float((fary >> sh) & 0xffu), ay, by);
    const float tnz = fmaf(float((nearz >> sh) & 0xffu), az, bz0), tfz = fmaf(float((farz >> sh) & 0xffu), az, bz0);
    float tnear = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
    float tfb = fminf(fminf(tfx, tfy), fminf(tfz, r.tfar));
    // slack for the rounding of the slab arithmetic
    tnear = fmaf(-fabsf(tnear), 4.8e-7f, tnear);
    tfb = fmaf(fabsf(tfb), 4.8e-7f, tfb);
    m[i] = cc[i] >= 0 && tnear <= tfb;
    tn[i] = tnear;
  }
  // inner children in front of the current closest hit, sorted by a fixed
  // network: the nearest is visited next, the others pushed far first.  A
  // leaf hit below can only move tfar closer; entries then behind it are
  // skipped when popped.  Any-hit rays end at their first hit, whatever its
  // distance, so they take the children in node order (per 1080p
  // large-mesh frame, shadow rays 33.2 -> 31.9 ms).
  float d0 = kInf, d1 = kInf, d2 = kInf, d3 = kInf;
  uint32_t n0 = 0, n1 = 0, n2 = 0, n3 = 0;
  if (m[0] && cnt[0] == 0 && tn[0] <= r.tfar) { d0 = tn[0]; n0 = uint32_t(cc[0]); }
  if (m[1] && cnt[1] == 0 && tn[1] <= r.tfar) { d1 = tn[1]; n1 = uint32_t(cc[1]); }
  if (m[2] && cnt[2] == 0 && tn[2] <= r.tfar) { d2 = tn[2]; n2 = uint32_t(cc[2]); }
  if (m[3] && cnt[3] == 0 && tn[3] <= r.tfar) { d3 = tn[3]; n3 = uint32_t(cc[3]); }
  if constexpr (!ANY) {
    cswap(d0, n0, d1, n1);
    cswap(d2, n2, d3, n3);
    cswap(d0, n0, d2, n2);
    cswap(d1, n1, d3, n3);
    cswap(d1, n1, d2, n2);
  }
  if (d3 != kInf) st.push(n3, d3);
  if (d2 != kInf) st.push(n2, d2);
  if (d1 != kInf) st.push(n1, d1);
  // this lane's leaf primitives as one list (children in order)
  const uint32_t l0 = (m[0] && cnt[0] > 0) ? cnt[0] : 0u, l1 = (m[1] && cnt[1] > 0) ? cnt[1] : 0u;
  const uint32_t l2 = (m[2] && cnt[2] > 0) ? cnt[2] : 0u, l3 = (m[3] && cnt[3] > 0) ? cnt[3] : 0u;
  const uint32_t e0 = l0, e1 = e0 + l1, e2 = e1 + l2, total = e2 + l3;
  const uint32_t b0 = uint32_t(cc[0]), b1 = uint32_t(cc[1]) - e0, b2 = uint32_t(cc[2]) - e1, b3 = uint32_t(cc[3]) - e2;
  bool leafHit = false;
  if constexpr (!POOL) {
    // one lane on its own: its primitives in order, the same (t, primitive) minimum
    unsigned long long best = ~0ull;
    for (uint32_t k = 0; k < total; k++) {
      const uint32_t prim = k >= e2 ? b3 + k : k >= e1 ? b2 + k : k >= e0 ? b1 + k : b0 + k;
      float tf = r.tfar;
      if (isect_prim(s, prim, r.o, r.d, r.tmin, tf)) {
        const unsigned long long key = (uint64_t(__float_as_uint(tf)) << 32) | prim;
        if (key < best) best = key;
        if (ANY) break;
      }
    }
    if (best != ~0ull) {
      r.tfar = __uint_as_float(uint32_t(best >> 32));
      r.hitp = uint32_t(best);
      r.hit = true;
      leafHit = true;
    }
  } else {
  const uint32_t lane = __lane_id();
  const uint32_t pre =  // this lane's first pool index
      total ? __hip_atomic_fetch_add(pool.ctr, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) : 0u;
  wave_lds_fence();
  const uint32_t sum = *pool.ctr;  // wave-uniform
  if (sum > 0) {
    // Post the entries that fit the pool; a lane whose entries run past it
    // (the wave has more than kPoolSize, which is rare) tests those itself.
    const uint32_t cap = s.pool_cap;
    const uint32_t fit = pre < cap ? min(total, cap - pre) : 0u;
    unsigned long long best = ~0ull;
    for (uint32_t k = 0; k < total; k++) {
      const uint32_t prim = k >= e2 ? b3 + k : k >= e1 ? b2 + k : k >= e0 ? b1 + k : b0 + k;
      if (k < fit) {
        pool.prim[pre + k] = prim;
        pool.owner[pre + k] = lane;
      } else {
        float tf = r.tfar;
        if (isect_prim(s, prim, r.o, r.d, r.tmin, tf)) {
          const unsigned long long key = (uint64_t(__float_as_uint(tf)) << 32) | prim;
          if (key < best) best = key;  // unsigned order: (t, primitive)
        }
      }
    }
    if (total) pool.key[lane] = best;
    wave_lds_fence();
    const unsigned long long act = __ballot(true);
    const uint32_t nact = (uint32_t)__popcll(act);
    const uint32_t rank = (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
    const uint32_t cnt_r = min(cap, sum);
    for (uint32_t jb = 0; jb < cnt_r; jb += nact) {
      const uint32_t j = jb + rank;
      const bool valid = j < cnt_r;
      const uint32_t own = valid ? pool.owner[j] : lane;
      // the owner's ray, read from its registers (all lanes of the step take part)
      const f3 o = mk(__shfl(r.o.x, own), __shfl(r.o.y, own), __shfl(r.o.z, own));
      const f3 d = mk(__shfl(r.d.x, own), __shfl(r.d.y, own), __shfl(r.d.z, own));
      const float tmin = __shfl(r.tmin, own);
      float tfar = __shfl(r.tfar, own);
      if (valid) {
        const uint32_t prim = pool.prim[j];
        if (isect_prim(s, prim, o, d, tmin, tfar))
          __hip_atomic_fetch_min(&pool.key[own], (uint64_t(__float_as_uint(tfar)) << 32) | prim, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WAVEFRONT);
      }
    }
    wave_lds_fence();
    if (total) {
      const unsigned long long k = pool.key[lane];
      if (k != ~0ull) {
        r.tfar = __uint_as_float(uint32_t(k >> 32));
        r.hitp = uint32_t(k);
        r.hit = true;
        leafHit = true;
      }
    }
    if (rank == 0) *pool.ctr = 0u;  // ready for the next step
    wave_lds_fence();
  }
  }
  if (ANY && leafHit) return true;
  if (d0 != kInf && d0 <= r.tfar) {
    r.node = n0;
    return false;
  }
  // pop, skipping subtrees entered beyond the closest hit found since the push
  while (st.sp > 0) {
    uint32_t n;
    float d;
    st.pop(n, d);
    if (d <= r.tfar) {
      r.node = n;
      return false;
    }
  }
  return true;
}

// Stream traversal for kernels that own a batch of rays.  A lane whose ray
// finishes is refilled from the batch while the other lanes of its wave keep
// traversing.  Each wave takes rays from a private chunk of kChunk
// consecutive items and grabs the next chunk with one atomic on `fetch`, so
// the shared counter sees one atomic per kChunk rays (with fetch == nullptr
// each wave takes the 64 items at its own position once).
// Job: bool load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax, uint32_t& tag)
//      void finish(uint32_t i, uint32_t tag, const TraceRay& r)
//      static constexpr bool kTagIsIndex; with it the tag is the item index
//      i itself, and the stream keeps one register for both
//      static constexpr bool kSuspend; with it, a ray that has visited
//      job.budget nodes in this launch is offered to the job, which may keep
//      it for a later launch (per-ray costs are heavy-tailed: near-grazing
//      rays over dense meshes visit thousands of nodes, and a launch cannot
//      end before its slowest ray):
//      uint32_t budget;
//      void resume(uint32_t tag, TraceRay& r, TraceStack& st)   (after trace_begin)
//      bool suspend(uint32_t tag, const TraceRay& r, const TraceStack& st)
// Launch with kTraceBlock threads per block; every lane must call this.
template <bool ANY, class Job>
__device__ __forceinline__ void trace_stream(const DevScene& s, uint32_t n, uint32_t* fetch, Job& job) {
  __shared__ uint32_t lds_node[kLdsStack * kTraceBlock];
  __shared__ float lds_dist[kLdsStack * kTraceBlock];
  __shared__ uint32_t pool_ctr[kTraceBlock / 64], pool_prim[kTraceBlock / 64][kPoolSize],
      pool_owner[kTraceBlock / 64][kPoolSize];
  __shared__ unsigned long long pool_key[kTraceBlock];
  constexpr int kRefillBelow = 48;  // refill once fewer lanes than this are busy
  constexpr uint32_t kChunk = 128;
  const uint32_t lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  TraceOverflow ov;
  TraceStack st;
  st.node = (lds_u32*)(lds_node + threadIdx.x);
  st.dist = (lds_f32*)(lds_dist + threadIdx.x);
  st.ov_node = ov.node;
  st.ov_dist = ov.dist;
  const uint32_t wv = threadIdx.x >> 6;
  const PrimPool pool{(lds_u32*)&pool_ctr[wv], (lds_u32*)pool_prim[wv], (lds_u32*)pool_owner[wv],
                      (__attribute__((address_space(3))) unsigned long long*)&pool_key[wv * 64]};
  if (lane == 0) *pool.ctr = 0u;
  wave_lds_fence();
  TraceRay r;
  bool active = false, exhausted = false;
  uint32_t idx = 0, tag = 0, visits = 0;
  uint32_t cbase = 0, cleft = 0;  // the wave's chunk (wave-uniform)
  for (;;) {
    if (!exhausted || cleft > 0) {
      unsigned long long need = __ballot(!active);
      if (need) {
        const uint32_t want = (uint32_t)__popcll(need);
        const uint32_t rank = (uint32_t)__popcll(need & below);
        uint32_t i = 0xffffffffu;
        if (fetch) {
          uint32_t nbase = 0, nleft = 0;
          if (cleft < want && !exhausted) {
            uint32_t b = 0;
            if (lane == 0) b = atomicAdd(fetch, kChunk);
            b = __shfl(b, 0);
            if (b >= n) {
              exhausted = true;
            } else {
              nbase = b;
              nleft = min(kChunk, n - b);
              if (b + kChunk >= n) exhausted = true;
            }
          }
          if (rank < cleft)
            i = cbase + rank;
          else if (rank - cleft < nleft)
            i = nbase + (rank - cleft);
          const uint32_t fromOld = min(want, cleft);
          cbase += fromOld;
          cleft -= fromOld;
          if (nleft) {
            const uint32_t fromNew = min(want - fromOld, nleft);
            cbase = nbase + fromNew;
            cleft = nleft - fromNew;
          }
        } else {
          i = blockIdx.x * kTraceBlock + (threadIdx.x & ~63u) + lane;
          exhausted = true;
        }
        if (!active && i < n) {
          f3 o, d;
          float tmin, tmax;
          if (job.load(i, o, d, tmin, tmax, tag)) {
            if constexpr (!Job::kTagIsIndex) idx = i;
            trace_begin(r, st, o, d, tmin, tmax);
            if constexpr (Job::kSuspend) job.resume(tag, r, st);
            visits = 0;
            active = true;
          }
        }
      }
    }
    if (__ballot(active) == 0ull) {
      if (exhausted && cleft == 0) break;
      continue;
    }
    for (int k = 0; k < 256; k++) {
      if (active) {
        if (trace_step<ANY, true>(s, r, st, pool)) {
          if (!ANY) trace_finish(s, r);
          job.finish(Job::kTagIsIndex ? tag : idx, tag, r);
          active = false;
        } else if constexpr (Job::kSuspend) {
          if (++visits >= job.budget) {
            if (job.suspend(tag, r, st)) active = false;
            visits = 0;
          }
        }
      }
      unsigned long long busy = __ballot(active);
      if (busy == 0ull) break;
      if ((!exhausted || cleft > 0) && __popcll(busy) < kRefillBelow) break;
    }
  }
}

// Single-ray traversal for kernels that trace a data-dependent number of rays
// per lane (launched with at most kTraceBlock threads per block).
template <bool ANY>
__device__ __forceinline__ bool trace(const DevScene& s, f3 o, f3 d, float tmin, float tmax, Hit& h) {
  __shared__ uint32_t lds_node1[kLdsStack * kTraceBlock];
  __shared__ float lds_dist1[kLdsStack * kTraceBlock];

  TraceRay r;
  TraceOverflow ov;
  TraceStack st;
  st.node = (lds_u32*)(lds_node1 + threadIdx.x);
  st.dist = (lds_f32*)(lds_dist1 + threadIdx.x);
  st.ov_node = ov.node;
  st.ov_dist = ov.dist;
  trace_begin(r, st, o, d, tmin, tmax);
  // lanes loop independently here, so primitives are tested per lane
  while (!trace_step<ANY, false>(s, r, st, PrimPool{})) {
  }
  if (!ANY) trace_finish(s, r);
  h = r.h;
  return r.hit;
}

}  // namespace rad
