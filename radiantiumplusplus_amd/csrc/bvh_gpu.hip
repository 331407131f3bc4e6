// Device BVH build (replaces the Embree build, embree.cpp:18-43).
//
// 1. Centroids are ordered along a 63-bit Morton curve (hipCUB radix sort):
//    the starting order, so neighbouring lanes hold nearby primitives.
// 2. Top-down, one level of nodes at a time over the whole GPU.  A node
//    with more than kSmall primitives is split by binned SAH: 16 centroid
//    bins on each axis, accumulated by every primitive of the node
//    (wave-aggregated float atomics), the best axis and bin boundary chosen
//    by the surface-area heuristic, and the node's range partitioned stably
//    in place through a prefix sum of the go-left flags.  Nodes of at most
//    kSmall primitives are built to the leaves by one wave each, with an
//    exact SAH sweep over the centroid order of every axis.  Ranges of at
//    most max_leaf primitives become leaves.
// 3. The binary tree is collapsed into the 4-wide, one-cache-line nodes the
//    traversal reads (repeatedly opening the inner child of largest area),
//    again level by level.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <stdexcept>

#include "bvh_gpu.h"

namespace rad {
namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kBins = 16;
constexpr uint32_t kSmall = 64;

struct N2 {
  float lo[3];
  uint32_t left;   // inner: first of two consecutive children; leaf: first sorted primitive
  float hi[3];
  uint32_t count;  // 0: inner
};

__device__ __forceinline__ void atomic_min_f(float* a, float v) {
  if (v >= 0.f)
    atomicMin(reinterpret_cast<int*>(a), __float_as_int(v));
  else
    atomicMax(reinterpret_cast<unsigned int*>(a), __float_as_uint(v));
}
__device__ __forceinline__ void atomic_max_f(float* a, float v) {
  if (v >= 0.f)
    atomicMax(reinterpret_cast<int*>(a), __float_as_int(v));
  else
    atomicMin(reinterpret_cast<unsigned int*>(a), __float_as_uint(v));
}

__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

__device__ __forceinline__ float half_area(const float lo[3], const float hi[3]) {
  float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  if (!(dx >= 0.f && dy >= 0.f && dz >= 0.f)) return 0.f;
  return dx * dy + dy * dz + dz * dx;
}

// ---------------------------------------------------------------- ordering
__global__ void k_centroid_bounds(const float* b, uint32_t n, float* cb) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float c[3], lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    c[a] = i < n ? 0.5f * (b[6 * size_t(i) + a] + b[6 * size_t(i) + 3 + a]) : 0.f;
    lo[a] = wave_min(i < n ? c[a] : kInf);
    hi[a] = wave_max(i < n ? c[a] : -kInf);
  }
  if (__lane_id() == 0)
    for (int a = 0; a < 3; a++) {
      atomic_min_f(&cb[a], lo[a]);
      atomic_max_f(&cb[3 + a], hi[a]);
    }
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

__global__ void k_morton(const float* b, uint32_t n, const float* cb, uint64_t* keys, uint32_t* ids) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t q[3];
  for (int a = 0; a < 3; a++) {
    float c = 0.5f * (b[6 * size_t(i) + a] + b[6 * size_t(i) + 3 + a]);
    float e = cb[3 + a] - cb[a];
    float t = e > 0.f ? (c - cb[a]) / e : 0.5f;
    q[a] = uint32_t(fminf(fmaxf(t, 0.f), 1.f) * 2097151.f);
  }
  keys[i] = spread21(q[0]) | spread21(q[1]) << 1 | spread21(q[2]) << 2;
  ids[i] = i;
}

// sorted SoA bounds: lo.xyz + pad, hi.xyz + pad
__global__ void k_gather_bounds(const float* b, const uint32_t* ids, uint32_t n, float4* blo, float4* bhi) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* s = b + 6 * size_t(ids[i]);
  blo[i] = make_float4(s[0], s[1], s[2], 0.f);
  bhi[i] = make_float4(s[3], s[4], s[5], 0.f);
}

// ---------------------------------------------------------------- top-down build
// Large nodes (more than kSmall primitives) are split one level at a time
// over the whole GPU: their primitives add their boxes to 16 centroid bins
// on each axis, the surface-area heuristic picks the axis and bin boundary,
// and a flag scan partitions the node's range in place (stable, ping-pong
// buffers).  Nodes of at most kSmall primitives are finished by one wave
// each with an exact SAH sweep over every centroid on every axis.
struct Task {
  uint32_t node, begin, end, nl;  // nl: primitives going left (set by the split)
  float clo[3];
  uint32_t axis;                  // split axis; 3: median split of the current order
  float chi[3];
  uint32_t kbin;                  // first bin of the right child
  uint32_t child_task[2];         // next-level task of each child, kNone if it is a leaf or small
  uint32_t pad[2];
};

struct SmallTask {
  uint32_t node, begin, end, pad;
};

struct Bin {
  float lo[3];
  uint32_t count;
  float hi[3];
  uint32_t pad;
};

struct BuildCtr {
  uint32_t nodes;   // binary nodes allocated
  uint32_t tasks;   // large tasks of the next level
  uint32_t smalls;  // small tasks (all levels)
  uint32_t error;   // set when a capacity would be exceeded (nothing is written then)
};

struct Caps {
  uint32_t nodes, tasks, smalls;
};

__device__ __forceinline__ int bin_of(float c, float lo, float hi) {
  const float e = hi - lo;
  if (!(e > 0.f)) return 0;
  const int k = int((c - lo) * (float(kBins) / e));
  return min(max(k, 0), kBins - 1);
}

__device__ __forceinline__ float4 centroid(float4 lo, float4 hi) {
  return make_float4(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z), 0.f);
}
__device__ __forceinline__ float comp(float4 v, int a) { return a == 0 ? v.x : a == 1 ? v.y : v.z; }

__global__ void k_bins_clear(Bin* bins, uint32_t m) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  Bin e;
  for (int a = 0; a < 3; a++) {
    e.lo[a] = kInf;
    e.hi[a] = -kInf;
  }
  e.count = 0;
  e.pad = 0;
  bins[i] = e;
}

// Adds this lane's box to bins[slot]; lanes of the wave that share a slot
// are combined first (one set of atomics per distinct slot).
__device__ __forceinline__ void bin_add(Bin* bins, uint32_t slot, float4 lo, float4 hi) {
  uint64_t pending = __ballot(slot != kNone);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t target = __shfl(slot, leader);
    const bool mine = slot == target;
    const uint64_t group = __ballot(mine);
    pending &= ~group;
    float l0 = wave_min(mine ? lo.x : kInf), l1 = wave_min(mine ? lo.y : kInf), l2 = wave_min(mine ? lo.z : kInf);
    float h0 = wave_max(mine ? hi.x : -kInf), h1 = wave_max(mine ? hi.y : -kInf), h2 = wave_max(mine ? hi.z : -kInf);
    if (__lane_id() == leader) {
      Bin* e = &bins[target];
      atomic_min_f(&e->lo[0], l0);
      atomic_min_f(&e->lo[1], l1);
      atomic_min_f(&e->lo[2], l2);
      atomic_max_f(&e->hi[0], h0);
      atomic_max_f(&e->hi[1], h1);
      atomic_max_f(&e->hi[2], h2);
      atomicAdd(&e->count, uint32_t(__popcll(group)));
    }
  }
}

// Every primitive of a large node adds its box to one bin per axis.
__global__ void k_bins_accum(uint32_t n, const uint32_t* owner, const Task* tasks, const float4* blo, const float4* bhi,
                             Bin* bins) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t t = i < n ? owner[i] : kNone;
  float4 lo = make_float4(kInf, kInf, kInf, 0.f), hi = make_float4(-kInf, -kInf, -kInf, 0.f);
  int k[3] = {0, 0, 0};
  if (t != kNone) {
    const Task& tk = tasks[t];
    lo = blo[i];
    hi = bhi[i];
    const float4 c = centroid(lo, hi);
    k[0] = bin_of(c.x, tk.clo[0], tk.chi[0]);
    k[1] = bin_of(c.y, tk.clo[1], tk.chi[1]);
    k[2] = bin_of(c.z, tk.clo[2], tk.chi[2]);
  }
  for (int a = 0; a < 3; a++) bin_add(bins, t == kNone ? kNone : (t * 3 + a) * kBins + k[a], lo, hi);
}

// Centroid bounds of the next level's tasks.
__global__ void k_task_cbounds(uint32_t n, const uint32_t* owner, const float4* blo, const float4* bhi, Task* tasks) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t t = i < n ? owner[i] : kNone;
  float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
  if (t != kNone) c = centroid(blo[i], bhi[i]);
  uint64_t pending = __ballot(t != kNone);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t target = __shfl(t, leader);
    const bool mine = t == target;
    pending &= ~__ballot(mine);
    float l0 = wave_min(mine ? c.x : kInf), l1 = wave_min(mine ? c.y : kInf), l2 = wave_min(mine ? c.z : kInf);
    float h0 = wave_max(mine ? c.x : -kInf), h1 = wave_max(mine ? c.y : -kInf), h2 = wave_max(mine ? c.z : -kInf);
    if (__lane_id() == leader) {
      Task* e = &tasks[target];
      atomic_min_f(&e->clo[0], l0);
      atomic_min_f(&e->clo[1], l1);
      atomic_min_f(&e->clo[2], l2);
      atomic_max_f(&e->chi[0], h0);
      atomic_max_f(&e->chi[1], h1);
      atomic_max_f(&e->chi[2], h2);
    }
  }
}

__device__ __forceinline__ void grow(float lo[3], float hi[3], const Bin& b) {
  for (int a = 0; a < 3; a++) lo[a] = fminf(lo[a], b.lo[a]), hi[a] = fmaxf(hi[a], b.hi[a]);
}

// One thread per large task: best binned SAH split over the three axes,
// children allocated; a child of one primitive becomes a leaf, one of at
// most kSmall a small task, a larger one a task of the next level.
__global__ void k_split(Task* tasks, uint32_t nTasks, const Bin* bins, N2* n2, Task* next, SmallTask* smalls,
                        BuildCtr* ctr, uint32_t maxLeaf, Caps caps) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nTasks) return;
  Task tk = tasks[t];
  const uint32_t cnt = tk.end - tk.begin;
  float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
  for (int k = 0; k < kBins; k++) grow(lo, hi, bins[(t * 3) * kBins + k]);
  float best = kInf;
  uint32_t axis = 3, kbin = 0, nl = cnt / 2;
  float clo[2][3], chi[2][3];  // children boxes
  for (int c = 0; c < 2; c++)
    for (int a = 0; a < 3; a++) clo[c][a] = lo[a], chi[c][a] = hi[a];
  for (int a = 0; a < 3; a++) {
    const Bin* ba = bins + (t * 3 + a) * kBins;
    float sA[kBins];
    uint32_t sN[kBins];
    float slo[3] = {kInf, kInf, kInf}, shi[3] = {-kInf, -kInf, -kInf};
    uint32_t sc = 0;
    for (int k = kBins - 1; k >= 1; k--) {
      const Bin b = ba[k];
      grow(slo, shi, b);
      sc += b.count;
      sA[k] = half_area(slo, shi);
      sN[k] = sc;
    }
    float plo[3] = {kInf, kInf, kInf}, phi[3] = {-kInf, -kInf, -kInf};
    uint32_t pc = 0;
    for (int k = 1; k < kBins; k++) {
      const Bin b = ba[k - 1];
      grow(plo, phi, b);
      pc += b.count;
      if (pc == 0 || sN[k] == 0) continue;
      const float c = half_area(plo, phi) * float(pc) + sA[k] * float(sN[k]);
      if (c < best) {
        best = c;
        axis = uint32_t(a);
        kbin = uint32_t(k);
        nl = pc;
      }
    }
  }
  if (axis < 3) {  // exact children boxes from the chosen axis' bins
    const Bin* ba = bins + (t * 3 + axis) * kBins;
    for (int c = 0; c < 2; c++)
      for (int a = 0; a < 3; a++) clo[c][a] = kInf, chi[c][a] = -kInf;
    for (int k = 0; k < kBins; k++) grow(clo[k < int(kbin) ? 0 : 1], chi[k < int(kbin) ? 0 : 1], ba[k]);
  }
  const uint32_t child = atomicAdd(&ctr->nodes, 2u);
  if (child + 2 > caps.nodes) {
    atomicOr(&ctr->error, 1u);
    tk.axis = 3;
    tk.nl = cnt / 2;
    tk.child_task[0] = tk.child_task[1] = kNone;
    tasks[t] = tk;
    return;
  }
  N2 me;
  for (int a = 0; a < 3; a++) me.lo[a] = lo[a], me.hi[a] = hi[a];
  me.left = child;
  me.count = 0;
  n2[tk.node] = me;
  const uint32_t cb[2] = {tk.begin, tk.begin + nl}, ce[2] = {tk.begin + nl, tk.end};
  for (int s = 0; s < 2; s++) {
    const uint32_t m = ce[s] - cb[s];
    tk.child_task[s] = kNone;
    if (m == 1) {
      N2 leaf;
      for (int a = 0; a < 3; a++) leaf.lo[a] = clo[s][a], leaf.hi[a] = chi[s][a];
      leaf.left = cb[s];
      leaf.count = m;
      n2[child + s] = leaf;
    } else if (m <= kSmall) {  // the wave builder also decides leaves of at most max_leaf
      const uint32_t j = atomicAdd(&ctr->smalls, 1u);
      if (j >= caps.smalls) {
        atomicOr(&ctr->error, 2u);
        continue;
      }
      smalls[j] = SmallTask{child + s, cb[s], ce[s], 0};
    } else {
      const uint32_t j = atomicAdd(&ctr->tasks, 1u);
      if (j >= caps.tasks) {
        atomicOr(&ctr->error, 4u);
        continue;
      }
      Task c;
      c.node = child + s;
      c.begin = cb[s];
      c.end = ce[s];
      c.nl = 0;
      c.axis = 3;
      c.kbin = 0;
      for (int a = 0; a < 3; a++) c.clo[a] = kInf, c.chi[a] = -kInf;
      c.child_task[0] = c.child_task[1] = kNone;
      c.pad[0] = c.pad[1] = 0;
      next[j] = c;
      tk.child_task[s] = j;
    }
  }
  tk.axis = axis;
  tk.kbin = kbin;
  tk.nl = nl;
  tasks[t] = tk;
}

__device__ __forceinline__ bool goes_left(const Task& tk, uint32_t i, float4 lo, float4 hi) {
  if (tk.axis >= 3) return i - tk.begin < tk.nl;
  const float4 c = centroid(lo, hi);
  return bin_of(comp(c, int(tk.axis)), tk.clo[tk.axis], tk.chi[tk.axis]) < int(tk.kbin);
}

__global__ void k_flags(uint32_t n, const uint32_t* owner, const Task* tasks, const float4* blo, const float4* bhi,
                        uint32_t* flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = owner[i];
  flags[i] = t != kNone && goes_left(tasks[t], i, blo[i], bhi[i]) ? 1u : 0u;
}

// Stable partition of every large task's range; primitives of finished
// ranges keep their place.  scan: exclusive prefix sum of the flags.
__global__ void k_scatter(uint32_t n, const uint32_t* owner, const Task* tasks, const uint32_t* flags,
                          const uint32_t* scan, const uint32_t* ids, const float4* blo, const float4* bhi, uint32_t* ids2,
                          float4* blo2, float4* bhi2, uint32_t* owner2, BuildCtr* ctr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = owner[i];
  uint32_t p = i, o = kNone;
  if (t != kNone) {
    const Task& tk = tasks[t];
    const uint32_t before = scan[i] - scan[tk.begin];  // left-going primitives of the task before i
    if (flags[i]) {
      p = tk.begin + before;
      o = tk.child_task[0];
    } else {
      p = tk.begin + tk.nl + (i - tk.begin - before);
      o = tk.child_task[1];
    }
    if (p < tk.begin || p >= tk.end) {  // the bin counts and the flags disagree: report, write nothing
      atomicOr(&ctr->error, 8u);
      return;
    }
  }
  ids2[p] = ids[i];
  blo2[p] = blo[i];
  bhi2[p] = bhi[i];
  owner2[p] = o;
}

// One wave per small task (2 to kSmall = 64 primitives, one per lane): the
// whole subtree, splitting each range at the best SAH position over the
// centroid order of every axis (ties broken by position), with the range
// partitioned stably through LDS.  Ranges of at most max_leaf primitives
// become leaves unless the split is cheaper (the host builder's rule).
constexpr int kSmallWaves = 4;
__global__ void __launch_bounds__(64 * kSmallWaves) k_small(const SmallTask* smalls, uint32_t nSmall, uint32_t* ids,
                                                            float4* blo, float4* bhi, N2* n2, BuildCtr* ctr,
                                                            uint32_t maxLeaf, Caps caps) {
  __shared__ float4 sLo[kSmallWaves][kSmall], sHi[kSmallWaves][kSmall];
  __shared__ uint32_t sId[kSmallWaves][kSmall];
  __shared__ uint32_t sStk[kSmallWaves][kSmall][2];  // (node, begin | end << 8)
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  float4* Lo = sLo[wave];
  float4* Hi = sHi[wave];
  uint32_t* Id = sId[wave];
  for (uint32_t t = blockIdx.x * kSmallWaves + wave; t < nSmall; t += gridDim.x * kSmallWaves) {
    const SmallTask st = smalls[t];
    const int m = int(st.end - st.begin);
    if (lane < m) {
      Lo[lane] = blo[st.begin + lane];
      Hi[lane] = bhi[st.begin + lane];
      Id[lane] = ids[st.begin + lane];
    }
    int sp = 0;
    if (lane == 0) {
      sStk[wave][0][0] = st.node;
      sStk[wave][0][1] = uint32_t(m) << 8;
    }
    sp = 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    while (sp > 0) {
      sp--;
      const uint32_t node = __builtin_amdgcn_readfirstlane(sStk[wave][sp][0]);
      const uint32_t be = __builtin_amdgcn_readfirstlane(sStk[wave][sp][1]);
      const int b = int(be & 0xffu), e = int(be >> 8), cnt = e - b;
      const bool act = lane >= b && lane < e;
      const float4 mlo = act ? Lo[lane] : make_float4(kInf, kInf, kInf, 0.f);
      const float4 mhi = act ? Hi[lane] : make_float4(-kInf, -kInf, -kInf, 0.f);
      float nlo[3] = {wave_min(mlo.x), wave_min(mlo.y), wave_min(mlo.z)};
      float nhi[3] = {wave_max(mhi.x), wave_max(mhi.y), wave_max(mhi.z)};
      // candidate: this lane's primitive is the first of the right side in
      // (centroid[a], position) order
      const float4 mc = centroid(mlo, mhi);
      float best = kInf;
      int bestAxis = 0;
      for (int a = 0; a < 3; a++) {
        const float key = comp(mc, a);
        float llo[3] = {kInf, kInf, kInf}, lhi[3] = {-kInf, -kInf, -kInf};
        float rlo[3] = {kInf, kInf, kInf}, rhi[3] = {-kInf, -kInf, -kInf};
        int nl = 0;
        for (int k = b; k < e; k++) {
          const float4 l = Lo[k], h = Hi[k];
          const float ck = comp(centroid(l, h), a);
          const bool left = ck < key || (ck == key && k < lane);
          float* xl = left ? llo : rlo;
          float* xh = left ? lhi : rhi;
          xl[0] = fminf(xl[0], l.x), xl[1] = fminf(xl[1], l.y), xl[2] = fminf(xl[2], l.z);
          xh[0] = fmaxf(xh[0], h.x), xh[1] = fmaxf(xh[1], h.y), xh[2] = fmaxf(xh[2], h.z);
          nl += left ? 1 : 0;
        }
        if (act && nl > 0) {
          const float c = half_area(llo, lhi) * float(nl) + half_area(rlo, rhi) * float(cnt - nl);
          if (c < best) {
            best = c;
            bestAxis = a;
          }
        }
      }
      // wave argmin (lowest lane among equal costs)
      float wc = best;
      int wl = act && best < kInf ? lane : 64;
      for (int o = 32; o > 0; o >>= 1) {
        const float oc = __shfl_xor(wc, o);
        const int ol = __shfl_xor(wl, o);
        if (oc < wc || (oc == wc && ol < wl)) wc = oc, wl = ol;
      }
      // a range of at most max_leaf primitives stays a leaf unless splitting
      // is cheaper: leaf cost n against 1 + SAH cost / parent area
      if (uint32_t(cnt) <= maxLeaf) {
        const float A = half_area(nlo, nhi);
        const float splitCost = A > 0.f && wl < 64 ? 1.f + wc / A : kInf;
        if (cnt <= 1 || float(cnt) <= splitCost) {
          if (lane == 0) {
            N2 leaf;
            for (int a = 0; a < 3; a++) leaf.lo[a] = nlo[a], leaf.hi[a] = nhi[a];
            leaf.left = st.begin + uint32_t(b);
            leaf.count = uint32_t(cnt);
            n2[node] = leaf;
          }
          continue;
        }
      }
      const int winner = wl;
      int axis = 0;
      float wkey = 0.f;
      const bool split_by_key = winner < 64;
      if (split_by_key) {
        axis = __shfl(bestAxis, winner);
        wkey = __shfl(comp(mc, axis), winner);
      }
      // stable partition of [b, e); the left size is counted from the
      // partition itself, so the positions always permute [b, e)
      const float myKey = comp(mc, axis);
      bool left = act && (split_by_key ? (myKey < wkey || (myKey == wkey && lane < winner)) : lane - b < cnt / 2);
      uint64_t lmask = __ballot(left);
      int nlw = __popcll(lmask);
      if (nlw == 0 || nlw == cnt) {  // no usable key split (e.g. NaN centroids): halve the range
        left = act && lane - b < cnt / 2;
        lmask = __ballot(left);
        nlw = __popcll(lmask);
      }
      const uint64_t below = (1ull << lane) - 1ull;
      const uint64_t rmask = __ballot(act && !left);
      const int pos = left ? b + __popcll(lmask & below) : b + nlw + __popcll(rmask & below);
      const uint32_t myId = act ? Id[lane] : 0u;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (act) {
        Lo[pos] = mlo;
        Hi[pos] = mhi;
        Id[pos] = myId;
      }
      uint32_t child = 0;
      if (lane == 0) child = atomicAdd(&ctr->nodes, 2u);
      child = __shfl(child, 0);
      if (child + 2 > caps.nodes) {
        if (lane == 0) atomicOr(&ctr->error, 1u);
        break;
      }
      if (lane == 0) {
        N2 me;
        for (int a = 0; a < 3; a++) me.lo[a] = nlo[a], me.hi[a] = nhi[a];
        me.left = child;
        me.count = 0;
        n2[node] = me;
        sStk[wave][sp][0] = child;
        sStk[wave][sp][1] = uint32_t(b) | uint32_t(b + nlw) << 8;
        sStk[wave][sp + 1][0] = child + 1;
        sStk[wave][sp + 1][1] = uint32_t(b + nlw) | uint32_t(e) << 8;
      }
      sp += 2;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (lane < m) {
      blo[st.begin + lane] = Lo[lane];
      bhi[st.begin + lane] = Hi[lane];
      ids[st.begin + lane] = Id[lane];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

__device__ inline void range_bounds(const float4* blo, const float4* bhi, uint32_t b, uint32_t e, float lo[3], float hi[3]) {
  lo[0] = lo[1] = lo[2] = kInf;
  hi[0] = hi[1] = hi[2] = -kInf;
  for (uint32_t i = b; i < e; i++) {
    float4 l = blo[i], h = bhi[i];
    lo[0] = fminf(lo[0], l.x), lo[1] = fminf(lo[1], l.y), lo[2] = fminf(lo[2], l.z);
    hi[0] = fmaxf(hi[0], h.x), hi[1] = fmaxf(hi[1], h.y), hi[2] = fmaxf(hi[2], h.z);
  }
}

__global__ void k_root_leaf(const float4* blo, const float4* bhi, uint32_t n, N2* n2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  N2 r;
  range_bounds(blo, bhi, 0, n, r.lo, r.hi);
  r.left = 0;
  r.count = n;
  n2[0] = r;
}

// ---------------------------------------------------------------- collapse to 4-wide nodes
struct Job {
  uint32_t n2, n4;
};

__device__ __forceinline__ float widen(float v, bool up) {
  float m = fabsf(v) * 4.0f * 1.1920929e-7f + 1e-30f;
  return up ? v + m : v - m;
}

__global__ void k_collapse(const Job* jobs, uint32_t nJobs, const N2* n2, BvhNode4* out, Job* next, uint32_t* ctr) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nJobs) return;
  Job jb = jobs[j];
  uint32_t kid[4];
  int nk = 0;
  const N2 x = n2[jb.n2];
  if (x.count > 0) {
    kid[nk++] = jb.n2;
  } else {
    kid[nk++] = x.left;
    kid[nk++] = x.left + 1;
  }
  while (nk < 4) {
    int bi = -1;
    float ba = -1.f;
    for (int k = 0; k < nk; k++) {
      const N2& c = n2[kid[k]];
      if (c.count > 0) continue;
      float a = half_area(c.lo, c.hi);
      if (a > ba) {
        ba = a;
        bi = k;
      }
    }
    if (bi < 0) break;
    uint32_t l = n2[kid[bi]].left;
    kid[bi] = l;
    kid[nk++] = l + 1;
  }
  // the node's inner children get consecutive ids, so siblings share cache lines
  int inner = 0;
  for (int s = 0; s < nk; s++) inner += n2[kid[s]].count > 0 ? 0 : 1;
  uint32_t base = 0, qbase = 0;
  if (inner) {
    base = atomicAdd(&ctr[0], uint32_t(inner));
    qbase = atomicAdd(&ctr[1], uint32_t(inner));
  }
  BvhNode4 o;
  for (int s = 0; s < 4; s++) {
    if (s >= nk) {
      o.lox[s] = o.loy[s] = o.loz[s] = kInf;
      o.hix[s] = o.hiy[s] = o.hiz[s] = -kInf;
      o.child[s] = -1;
      o.count[s] = 0;
      continue;
    }
    const N2 c = n2[kid[s]];
    o.lox[s] = widen(c.lo[0], false);
    o.loy[s] = widen(c.lo[1], false);
    o.loz[s] = widen(c.lo[2], false);
    o.hix[s] = widen(c.hi[0], true);
    o.hiy[s] = widen(c.hi[1], true);
    o.hiz[s] = widen(c.hi[2], true);
    if (c.count > 0) {
      o.child[s] = int32_t(c.left);
      o.count[s] = c.count;
    } else {
      next[qbase] = Job{kid[s], base};
      o.child[s] = int32_t(base);
      o.count[s] = 0;
      base++;
      qbase++;
    }
  }
  out[jb.n4] = o;
}

// Worst-case traversal stack entries of the nodes [begin, end) of one level,
// from those of the next level (launched deepest level first).
__global__ void k_stack_need(const BvhNode4* nodes, uint32_t begin, uint32_t end, uint32_t* need) {
  const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= end) return;
  const BvhNode4& nd = nodes[i];
  uint32_t inner = 0, deepest = 0;
  for (int s = 0; s < 4; s++)
    if (nd.child[s] >= 0 && nd.count[s] == 0) {
      inner++;
      deepest = max(deepest, need[nd.child[s]]);
    }
  need[i] = inner ? inner - 1 + deepest : 0u;
}

__global__ void k_gather_prims(const float4* rec, const uint32_t* ids, uint32_t n, float4* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* s = rec + 3 * size_t(ids[i]);
  out[3 * size_t(i)] = s[0];
  out[3 * size_t(i) + 1] = s[1];
  out[3 * size_t(i) + 2] = s[2];
}

// SAH cost (traversal 1, intersection 1, relative to the root area).
__global__ void k_sah(const BvhNode4* nodes, uint32_t m, float rootA, double* cost) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float c = 0.f;
  if (i < m) {
    const BvhNode4& nd = nodes[i];
    for (int s = 0; s < 4; s++) {
      if (nd.child[s] < 0) continue;
      float lo[3] = {nd.lox[s], nd.loy[s], nd.loz[s]}, hi[3] = {nd.hix[s], nd.hiy[s], nd.hiz[s]};
      float a = 2.f * half_area(lo, hi) / rootA;
      c += nd.count[s] > 0 ? a * float(nd.count[s]) : a;
    }
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (__lane_id() == 0 && c != 0.f) atomicAdd(cost, double(c));
}

// ---------------------------------------------------------------- quantised copy
// Step 2^e per axis, e the smallest with 255 * 2^e >= the node's extent
// (raised further if rounding keeps a bound out of reach); bounds rounded
// outward and verified with the traversal's own dequantisation.
__device__ bool quantize_axis(const float* clo, const float* chi, const bool* valid, float o, float ext, uint32_t& biased,
                              uint32_t& qlo, uint32_t& qhi) {
  int e = -126;
  if (ext > 0.f) {
    frexpf(ext / 255.f, &e);  // ext / 255 = m 2^e, m in [0.5, 1)
    e = max(e, -126);
  }
  for (; e <= 127; e++) {
    const uint32_t b = uint32_t(e + 127);
    const float step = quant_step(b);
    bool ok = true;
    uint32_t lo = 0, hi = 0;
    for (int i = 0; i < 4 && ok; i++) {
      if (!valid[i]) continue;
      float fl = fminf(fmaxf(floorf((clo[i] - o) / step), 0.f), 255.f);
      float fh = fminf(fmaxf(ceilf((chi[i] - o) / step), 0.f), 255.f);
      uint32_t a = uint32_t(fl), c = uint32_t(fh);
      while (a > 0 && dequant(a, step, o) > clo[i]) a--;
      while (c < 255 && dequant(c, step, o) < chi[i]) c++;
      if (dequant(a, step, o) > clo[i] || dequant(c, step, o) < chi[i]) ok = false;
      lo |= a << (8 * i);
      hi |= c << (8 * i);
    }
    if (ok) {
      biased = b;
      qlo = lo;
      qhi = hi;
      return true;
    }
  }
  return false;
}

__global__ void k_quantize(const BvhNode4* in, uint32_t m, BvhNodeQ* out, uint32_t* error) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const BvhNode4 nd = in[i];
  bool valid[4];
  float o[3] = {kInf, kInf, kInf}, h[3] = {-kInf, -kInf, -kInf};
  for (int c = 0; c < 4; c++) {
    valid[c] = nd.child[c] >= 0;
    if (!valid[c]) continue;
    o[0] = fminf(o[0], nd.lox[c]), o[1] = fminf(o[1], nd.loy[c]), o[2] = fminf(o[2], nd.loz[c]);
    h[0] = fmaxf(h[0], nd.hix[c]), h[1] = fmaxf(h[1], nd.hiy[c]), h[2] = fmaxf(h[2], nd.hiz[c]);
  }
  BvhNodeQ q;
  uint32_t counts = 0;
  for (int c = 0; c < 4; c++) {
    q.child[c] = nd.child[c];
    counts |= (valid[c] ? min(nd.count[c], 255u) : 0u) << (8 * c);
  }
  q.counts = counts;
  q.pad = 0;
  if (!(valid[0] || valid[1] || valid[2] || valid[3])) {
    q.ox = q.oy = q.oz = 0.f;
    q.exps = 127u | 127u << 8 | 127u << 16;
    q.qlox = q.qhix = q.qloy = q.qhiy = q.qloz = q.qhiz = 0;
    out[i] = q;
    return;
  }
  uint32_t bx = 127, by = 127, bz = 127;
  bool ok = quantize_axis(nd.lox, nd.hix, valid, o[0], h[0] - o[0], bx, q.qlox, q.qhix) &&
            quantize_axis(nd.loy, nd.hiy, valid, o[1], h[1] - o[1], by, q.qloy, q.qhiy) &&
            quantize_axis(nd.loz, nd.hiz, valid, o[2], h[2] - o[2], bz, q.qloz, q.qhiz);
  if (!ok || nd.count[0] > 255u || nd.count[1] > 255u || nd.count[2] > 255u || nd.count[3] > 255u) atomicOr(error, 1u);
  q.ox = o[0];
  q.oy = o[1];
  q.oz = o[2];
  q.exps = bx | by << 8 | bz << 16;
  out[i] = q;
}

unsigned Grid(uint64_t n, unsigned t = 256) { return unsigned((n + t - 1) / t); }

}  // namespace

void QuantizeBvh4(const DeviceBuffer& nodes, uint32_t m, hipStream_t st, DeviceBuffer& out) {
  out.alloc(size_t(std::max(m, 1u)) * sizeof(BvhNodeQ));
  DeviceBuffer err;
  err.alloc(16);
  HipCheck(hipMemsetAsync(err.ptr, 0, 4, st), "quantize flag");
  if (m) hipLaunchKernelGGL(k_quantize, dim3(Grid(m)), dim3(256), 0, st, nodes.as<BvhNode4>(), m, out.as<BvhNodeQ>(),
                            err.as<uint32_t>());
  uint32_t e = 0;
  HipCheck(hipMemcpyAsync(&e, err.ptr, 4, hipMemcpyDeviceToHost, st), "quantize flag");
  HipCheck(hipStreamSynchronize(st), "quantize");
  if (e) throw std::runtime_error("BVH node bounds cannot be quantised (non-finite boxes?)");
}

void BuildBvh4Device(const float* bounds, const float4* records, uint32_t n, int maxLeaf, hipStream_t st,
                     DeviceBuffer& nodesOut, DeviceBuffer& primsOut, DeviceBvhStats& stats) {
  if (n == 0) throw std::runtime_error("scene has no primitives");
  if (maxLeaf < 1 || maxLeaf > 16) throw std::runtime_error("max_leaf_size must lie in [1, 16]");
  auto t0 = std::chrono::steady_clock::now();
  DeviceBuffer dB, dRec, dCb, dKeys, dKeys2, dIds, dIds2, dIds3, dLo, dHi, dLo2, dHi2, dOwn0, dOwn1, dTasks0, dTasks1,
      dSmall, dBins, dFlags, dScan, dN2, dCtr, dTmp;
  dB.alloc(size_t(n) * 24);
  dRec.alloc(size_t(n) * 48);
  HipCheck(hipMemcpyAsync(dB.ptr, bounds, size_t(n) * 24, hipMemcpyHostToDevice, st), "upload bounds");
  HipCheck(hipMemcpyAsync(dRec.ptr, records, size_t(n) * 48, hipMemcpyHostToDevice, st), "upload primitives");
  // ordering
  dCb.alloc(64);
  {
    float init[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
    HipCheck(hipMemcpyAsync(dCb.ptr, init, 24, hipMemcpyHostToDevice, st), "centroid init");
  }
  hipLaunchKernelGGL(k_centroid_bounds, dim3(Grid(n)), dim3(256), 0, st, dB.as<float>(), n, dCb.as<float>());
  dKeys.alloc(size_t(n) * 8);
  dKeys2.alloc(size_t(n) * 8);
  dIds.alloc(size_t(n) * 4);
  dIds2.alloc(size_t(n) * 4);
  hipLaunchKernelGGL(k_morton, dim3(Grid(n)), dim3(256), 0, st, dB.as<float>(), n, dCb.as<float>(), dKeys.as<uint64_t>(),
                     dIds.as<uint32_t>());
  {
    size_t tmpBytes = 0;
    HipCheck(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpBytes, dKeys.as<uint64_t>(), dKeys2.as<uint64_t>(),
                                                dIds.as<uint32_t>(), dIds2.as<uint32_t>(), int(n), 0, 63, st),
             "radix sort size");
    size_t scanBytes = 0;  // the level loop's flag scans reuse this scratch
    HipCheck(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, static_cast<uint32_t*>(nullptr),
                                              static_cast<uint32_t*>(nullptr), int(n), st),
             "scan size");
    dTmp.alloc(std::max<size_t>(std::max(tmpBytes, scanBytes), 16));
    HipCheck(hipcub::DeviceRadixSort::SortPairs(dTmp.ptr, tmpBytes, dKeys.as<uint64_t>(), dKeys2.as<uint64_t>(),
                                                dIds.as<uint32_t>(), dIds2.as<uint32_t>(), int(n), 0, 63, st),
             "radix sort");
  }
  dLo.alloc(size_t(n) * 16);
  dHi.alloc(size_t(n) * 16);
  hipLaunchKernelGGL(k_gather_bounds, dim3(Grid(n)), dim3(256), 0, st, dB.as<float>(), dIds2.as<uint32_t>(), n,
                     dLo.as<float4>(), dHi.as<float4>());
  // top-down levels
  const uint32_t maxNodes = 2 * n;
  dN2.alloc(size_t(maxNodes) * sizeof(N2));
  const uint32_t maxTasks = n / (kSmall + 1) + 2;
  const uint32_t maxSmall = n / 2 + 2;  // small tasks hold at least two primitives
  dTasks0.alloc(size_t(maxTasks) * sizeof(Task));
  dTasks1.alloc(size_t(maxTasks) * sizeof(Task));
  dSmall.alloc(size_t(maxSmall) * sizeof(SmallTask));
  dBins.alloc(size_t(maxTasks) * 3 * kBins * sizeof(Bin));
  dOwn0.alloc(size_t(n) * 4);
  dOwn1.alloc(size_t(n) * 4);
  dFlags.alloc(size_t(n) * 4 + 16);
  dScan.alloc(size_t(n) * 4 + 16);
  dLo2.alloc(size_t(n) * 16);
  dHi2.alloc(size_t(n) * 16);
  dIds3.alloc(size_t(n) * 4);
  dCtr.alloc(64);
  BuildCtr* ctr = dCtr.as<BuildCtr>();
  const Caps caps{maxNodes, maxTasks, maxSmall};
  uint32_t* ids = dIds2.as<uint32_t>();  // Morton order
  uint32_t* ids2 = dIds3.as<uint32_t>();
  float4* lo = dLo.as<float4>();
  float4* hi = dHi.as<float4>();
  float4* lo2 = dLo2.as<float4>();
  float4* hi2 = dHi2.as<float4>();
  uint32_t levels = 0, nSmall = 0;
  BuildCtr c0{1, 0, 0, 0};
  if (n == 1) {
    hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(64), 0, st, lo, hi, n, dN2.as<N2>());
  } else if (n <= kSmall) {
    SmallTask root{0, 0, n, 0};
    HipCheck(hipMemcpyAsync(dSmall.ptr, &root, sizeof(root), hipMemcpyHostToDevice, st), "root task");
    c0.smalls = 1;
    nSmall = 1;
  } else {
    Task root{};
    root.node = 0;
    root.begin = 0;
    root.end = n;
    root.axis = 3;
    root.child_task[0] = root.child_task[1] = kNone;
    HipCheck(hipMemcpyAsync(&root.clo, dCb.ptr, 12, hipMemcpyDeviceToHost, st), "centroid bounds");
    HipCheck(hipMemcpyAsync(&root.chi, dCb.as<float>() + 3, 12, hipMemcpyDeviceToHost, st), "centroid bounds");
    HipCheck(hipStreamSynchronize(st), "centroid bounds");
    HipCheck(hipMemcpyAsync(dTasks0.ptr, &root, sizeof(Task), hipMemcpyHostToDevice, st), "root task");
    HipCheck(hipMemsetAsync(dOwn0.ptr, 0, size_t(n) * 4, st), "owners");
  }
  HipCheck(hipMemcpyAsync(ctr, &c0, sizeof(c0), hipMemcpyHostToDevice, st), "counters");
  if (n > kSmall) {
    uint32_t nTasks = 1;
    Task* cur = dTasks0.as<Task>();
    Task* nxt = dTasks1.as<Task>();
    uint32_t* own = dOwn0.as<uint32_t>();
    uint32_t* own2 = dOwn1.as<uint32_t>();
    while (nTasks > 0) {
      if (++levels > 128) throw std::runtime_error("BVH build did not converge");
      hipLaunchKernelGGL(k_bins_clear, dim3(Grid(uint64_t(nTasks) * 3 * kBins)), dim3(256), 0, st, dBins.as<Bin>(),
                         nTasks * 3 * kBins);
      hipLaunchKernelGGL(k_bins_accum, dim3(Grid(n)), dim3(256), 0, st, n, own, cur, lo, hi, dBins.as<Bin>());
      HipCheck(hipMemsetAsync(&ctr->tasks, 0, 4, st), "level counter");
      hipLaunchKernelGGL(k_split, dim3(Grid(nTasks, 64)), dim3(64), 0, st, cur, nTasks, dBins.as<Bin>(), dN2.as<N2>(), nxt,
                         dSmall.as<SmallTask>(), ctr, uint32_t(maxLeaf), caps);
      hipLaunchKernelGGL(k_flags, dim3(Grid(n)), dim3(256), 0, st, n, own, cur, lo, hi, dFlags.as<uint32_t>());
      size_t scanBytes = dTmp.bytes;
      HipCheck(hipcub::DeviceScan::ExclusiveSum(dTmp.ptr, scanBytes, dFlags.as<uint32_t>(), dScan.as<uint32_t>(), int(n), st),
               "flag scan");
      hipLaunchKernelGGL(k_scatter, dim3(Grid(n)), dim3(256), 0, st, n, own, cur, dFlags.as<uint32_t>(),
                         dScan.as<uint32_t>(), ids, lo, hi, ids2, lo2, hi2, own2, ctr);
      std::swap(ids, ids2);
      std::swap(lo, lo2);
      std::swap(hi, hi2);
      std::swap(own, own2);
      BuildCtr h{};
      HipCheck(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st), "level readback");
      HipCheck(hipStreamSynchronize(st), "BVH level");
      if (h.error || h.nodes > maxNodes || h.tasks > maxTasks || h.smalls > maxSmall)
        throw std::runtime_error("BVH build exceeded its buffers");
      nTasks = h.tasks;
      nSmall = h.smalls;
      if (nTasks) hipLaunchKernelGGL(k_task_cbounds, dim3(Grid(n)), dim3(256), 0, st, n, own, lo, hi, nxt);
      std::swap(cur, nxt);
    }
  }
  if (nSmall) {
    const uint32_t blocks = std::min<uint32_t>((nSmall + kSmallWaves - 1) / kSmallWaves, 65535u * 4u);
    hipLaunchKernelGGL(k_small, dim3(blocks), dim3(64 * kSmallWaves), 0, st, dSmall.as<SmallTask>(), nSmall, ids, lo, hi,
                       dN2.as<N2>(), ctr, uint32_t(maxLeaf), caps);
    BuildCtr h{};
    HipCheck(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st), "small readback");
    HipCheck(hipStreamSynchronize(st), "BVH small subtrees");
    if (h.error || h.nodes > maxNodes) throw std::runtime_error("BVH build exceeded its buffers");
  }
  const uint32_t* sorted = ids;
  // collapse
  BuildCtr hc{};
  HipCheck(hipMemcpyAsync(&hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, st), "node count");
  HipCheck(hipStreamSynchronize(st), "BVH build");
  const uint32_t n2count = hc.nodes;  // the counter starts at 1 (the root)
  nodesOut.alloc(size_t(n2count) * sizeof(BvhNode4));
  DeviceBuffer dJobs0, dJobs1, dJc;
  dJobs0.alloc(size_t(n2count) * sizeof(Job));
  dJobs1.alloc(size_t(n2count) * sizeof(Job));
  dJc.alloc(16);
  Job rj{0, 0};
  HipCheck(hipMemcpyAsync(dJobs0.ptr, &rj, sizeof(rj), hipMemcpyHostToDevice, st), "root job");
  uint32_t jc[2] = {1, 0};
  uint32_t nJobs = 1, depth = 0;
  Job* jcur = dJobs0.as<Job>();
  Job* jnxt = dJobs1.as<Job>();
  std::vector<uint32_t> levelStart{0u};  // the nodes of level k are [levelStart[k], levelStart[k + 1])
  while (nJobs > 0) {
    depth++;
    levelStart.push_back(jc[0]);
    jc[1] = 0;
    HipCheck(hipMemcpyAsync(dJc.ptr, jc, 8, hipMemcpyHostToDevice, st), "collapse counters");
    hipLaunchKernelGGL(k_collapse, dim3(Grid(nJobs, 64)), dim3(64), 0, st, jcur, nJobs, dN2.as<N2>(),
                       nodesOut.as<BvhNode4>(), jnxt, dJc.as<uint32_t>());
    HipCheck(hipMemcpyAsync(jc, dJc.ptr, 8, hipMemcpyDeviceToHost, st), "collapse readback");
    HipCheck(hipStreamSynchronize(st), "collapse level");
    nJobs = jc[1];
    std::swap(jcur, jnxt);
  }
  const uint32_t n4 = jc[0];
  levelStart.push_back(n4);
  uint32_t stackNeed = 0;
  {
    DeviceBuffer dNeed;
    dNeed.alloc(size_t(n4) * sizeof(uint32_t));
    for (size_t k = levelStart.size() - 1; k-- > 0;) {
      const uint32_t b = levelStart[k], e = levelStart[k + 1];
      if (e > b) hipLaunchKernelGGL(k_stack_need, dim3(Grid(e - b)), dim3(256), 0, st, nodesOut.as<BvhNode4>(), b, e,
                                    dNeed.as<uint32_t>());
    }
    HipCheck(hipMemcpyAsync(&stackNeed, dNeed.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "stack need");
    HipCheck(hipStreamSynchronize(st), "stack need");
  }
  primsOut.alloc(size_t(n) * 48);
  hipLaunchKernelGGL(k_gather_prims, dim3(Grid(n)), dim3(256), 0, st, dRec.as<float4>(), sorted, n, primsOut.as<float4>());
  // statistics: SAH cost relative to the root box
  BvhNode4 root;
  HipCheck(hipMemcpyAsync(&root, nodesOut.ptr, sizeof(root), hipMemcpyDeviceToHost, st), "root readback");
  HipCheck(hipStreamSynchronize(st), "BVH finish");
  float rlo[3] = {kInf, kInf, kInf}, rhi[3] = {-kInf, -kInf, -kInf};
  for (int s = 0; s < 4; s++) {
    if (root.child[s] < 0) continue;
    rlo[0] = std::min(rlo[0], root.lox[s]), rlo[1] = std::min(rlo[1], root.loy[s]), rlo[2] = std::min(rlo[2], root.loz[s]);
    rhi[0] = std::max(rhi[0], root.hix[s]), rhi[1] = std::max(rhi[1], root.hiy[s]), rhi[2] = std::max(rhi[2], root.hiz[s]);
  }
  float dx = rhi[0] - rlo[0], dy = rhi[1] - rlo[1], dz = rhi[2] - rlo[2];
  float rootA = 2.f * (dx * dy + dy * dz + dz * dx);
  double sah = 1.0;
  if (rootA > 0.f) {
    DeviceBuffer dCost;
    dCost.alloc(16);
    HipCheck(hipMemsetAsync(dCost.ptr, 0, 8, st), "sah");
    hipLaunchKernelGGL(k_sah, dim3(Grid(n4)), dim3(256), 0, st, nodesOut.as<BvhNode4>(), n4, rootA, dCost.as<double>());
    double c = 0;
    HipCheck(hipMemcpyAsync(&c, dCost.ptr, 8, hipMemcpyDeviceToHost, st), "sah readback");
    HipCheck(hipStreamSynchronize(st), "sah");
    sah += c;
  }
  HipCheck(hipGetLastError(), "BVH kernels");
  stats.node_count = n4;
  stats.depth = depth;
  stats.stack_need = stackNeed;
  stats.levels = levels;
  stats.sah_cost = float(sah);
  stats.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace rad
