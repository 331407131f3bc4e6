// Device BVH build (replaces the Embree build, embree.cpp:18-43).
//
// 1. Centroids are ordered along a 63-bit Morton curve (hipCUB radix sort),
//    so every node owns a contiguous range of the sorted primitives and
//    splitting never moves data.
// 2. Top-down, one level of nodes at a time over the whole GPU.  A node
//    with more than kSmall primitives is split at the best of 15 binned
//    candidates: its range is cut into 16 equal bins whose bounds are
//    accumulated by every primitive (wave-aggregated float atomics through a
//    per-position owner array), and the surface-area heuristic picks the bin
//    boundary.  Smaller nodes sweep every split position exactly.
//    Ranges of at most max_leaf primitives become leaves.
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

struct Task {
  uint32_t node, begin, end;
  uint32_t bin;    // bin block (kBins entries) for large nodes, kNone otherwise
  uint32_t split;
  uint32_t child_task[2];
  uint32_t pad;
};

struct Bin {
  float lo[3];
  uint32_t count;
  float hi[3];
  uint32_t pad;
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

// Every primitive adds its box to its node's bin (large nodes only).
__global__ void k_bins_accum(uint32_t n, const uint32_t* owner, const Task* tasks, const float4* blo, const float4* bhi,
                             Bin* bins) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t slot = kNone;
  float4 lo = make_float4(kInf, kInf, kInf, 0.f), hi = make_float4(-kInf, -kInf, -kInf, 0.f);
  if (i < n) {
    uint32_t t = owner[i];
    if (t != kNone) {
      Task tk = tasks[t];
      if (tk.bin != kNone) {
        uint32_t cnt = tk.end - tk.begin;
        uint32_t k = uint32_t((uint64_t(i - tk.begin) * kBins) / cnt);
        slot = tk.bin * kBins + k;
        lo = blo[i];
        hi = bhi[i];
      }
    }
  }
  // one set of atomics per wave when the whole wave feeds one bin
  uint32_t first = __shfl(slot, 0);
  bool same = __ballot(slot == first) == ~0ull;
  if (same) {
    if (first == kNone) return;
    float l0 = wave_min(lo.x), l1 = wave_min(lo.y), l2 = wave_min(lo.z);
    float h0 = wave_max(hi.x), h1 = wave_max(hi.y), h2 = wave_max(hi.z);
    if (__lane_id() == 0) {
      Bin* e = &bins[first];
      atomic_min_f(&e->lo[0], l0);
      atomic_min_f(&e->lo[1], l1);
      atomic_min_f(&e->lo[2], l2);
      atomic_max_f(&e->hi[0], h0);
      atomic_max_f(&e->hi[1], h1);
      atomic_max_f(&e->hi[2], h2);
      atomicAdd(&e->count, 64u);
    }
    return;
  }
  if (slot == kNone) return;
  Bin* e = &bins[slot];
  atomic_min_f(&e->lo[0], lo.x);
  atomic_min_f(&e->lo[1], lo.y);
  atomic_min_f(&e->lo[2], lo.z);
  atomic_max_f(&e->hi[0], hi.x);
  atomic_max_f(&e->hi[1], hi.y);
  atomic_max_f(&e->hi[2], hi.z);
  atomicAdd(&e->count, 1u);
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

struct BuildCtr {
  uint32_t nodes;   // binary nodes allocated
  uint32_t tasks;   // tasks of the next level
  uint32_t bins;    // bin blocks of the next level
  uint32_t error;   // set when a capacity would be exceeded (nothing is written then)
};

struct Caps {
  uint32_t nodes, tasks, bins;
};

__global__ void k_split(Task* tasks, uint32_t nTasks, const Bin* bins, const float4* blo, const float4* bhi, N2* n2,
                        Task* next, BuildCtr* ctr, uint32_t maxLeaf, Caps caps) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nTasks) return;
  Task tk = tasks[t];
  const uint32_t cnt = tk.end - tk.begin;
  float lo[3], hi[3];
  uint32_t split;
  if (tk.bin != kNone) {
    Bin b[kBins];
    for (int k = 0; k < kBins; k++) b[k] = bins[tk.bin * kBins + k];
    // suffix areas
    float sA[kBins];
    float slo[3] = {kInf, kInf, kInf}, shi[3] = {-kInf, -kInf, -kInf};
    for (int k = kBins - 1; k >= 0; k--) {
      for (int a = 0; a < 3; a++) slo[a] = fminf(slo[a], b[k].lo[a]), shi[a] = fmaxf(shi[a], b[k].hi[a]);
      sA[k] = half_area(slo, shi);
    }
    for (int a = 0; a < 3; a++) lo[a] = slo[a], hi[a] = shi[a];
    float plo[3] = {kInf, kInf, kInf}, phi[3] = {-kInf, -kInf, -kInf};
    float best = kInf;
    split = tk.begin + cnt / 2;
    for (int k = 1; k < kBins; k++) {
      for (int a = 0; a < 3; a++) plo[a] = fminf(plo[a], b[k - 1].lo[a]), phi[a] = fmaxf(phi[a], b[k - 1].hi[a]);
      uint32_t pos = tk.begin + uint32_t((uint64_t(k) * cnt) / kBins);
      uint32_t nl = pos - tk.begin, nr = tk.end - pos;
      if (nl == 0 || nr == 0) continue;
      float c = half_area(plo, phi) * float(nl) + sA[k] * float(nr);
      if (c < best) {
        best = c;
        split = pos;
      }
    }
  } else {
    // exact sweep over every split position of a small range
    float sA[kSmall];
    float slo[3] = {kInf, kInf, kInf}, shi[3] = {-kInf, -kInf, -kInf};
    for (uint32_t j = cnt; j-- > 0;) {
      float4 l = blo[tk.begin + j], h = bhi[tk.begin + j];
      slo[0] = fminf(slo[0], l.x), slo[1] = fminf(slo[1], l.y), slo[2] = fminf(slo[2], l.z);
      shi[0] = fmaxf(shi[0], h.x), shi[1] = fmaxf(shi[1], h.y), shi[2] = fmaxf(shi[2], h.z);
      sA[j] = half_area(slo, shi);
    }
    for (int a = 0; a < 3; a++) lo[a] = slo[a], hi[a] = shi[a];
    float plo[3] = {kInf, kInf, kInf}, phi[3] = {-kInf, -kInf, -kInf};
    float best = kInf;
    split = tk.begin + cnt / 2;
    for (uint32_t j = 1; j < cnt; j++) {
      float4 l = blo[tk.begin + j - 1], h = bhi[tk.begin + j - 1];
      plo[0] = fminf(plo[0], l.x), plo[1] = fminf(plo[1], l.y), plo[2] = fminf(plo[2], l.z);
      phi[0] = fmaxf(phi[0], h.x), phi[1] = fmaxf(phi[1], h.y), phi[2] = fmaxf(phi[2], h.z);
      float c = half_area(plo, phi) * float(j) + sA[j] * float(cnt - j);
      if (c < best) {
        best = c;
        split = tk.begin + j;
      }
    }
  }
  const uint32_t child = atomicAdd(&ctr->nodes, 2u);
  if (child + 2 > caps.nodes) {
    atomicOr(&ctr->error, 1u);
    return;
  }
  N2 me;
  for (int a = 0; a < 3; a++) me.lo[a] = lo[a], me.hi[a] = hi[a];
  me.left = child;
  me.count = 0;
  n2[tk.node] = me;
  const uint32_t cb[2] = {tk.begin, split}, ce[2] = {split, tk.end};
  for (int s = 0; s < 2; s++) {
    uint32_t m = ce[s] - cb[s];
    if (m <= maxLeaf) {
      N2 leaf;
      range_bounds(blo, bhi, cb[s], ce[s], leaf.lo, leaf.hi);
      leaf.left = cb[s];
      leaf.count = m;
      n2[child + s] = leaf;
      tk.child_task[s] = kNone;
    } else {
      uint32_t j = atomicAdd(&ctr->tasks, 1u);
      uint32_t bin = m > kSmall ? atomicAdd(&ctr->bins, 1u) : kNone;
      if (j >= caps.tasks || (bin != kNone && bin >= caps.bins)) {
        atomicOr(&ctr->error, 2u);
        tk.child_task[s] = kNone;
        continue;
      }
      Task c;
      c.node = child + s;
      c.begin = cb[s];
      c.end = ce[s];
      c.bin = bin;
      c.split = 0;
      c.child_task[0] = c.child_task[1] = kNone;
      c.pad = 0;
      next[j] = c;
      tk.child_task[s] = j;
    }
  }
  tk.split = split;
  tasks[t] = tk;
}

__global__ void k_owner(uint32_t n, const uint32_t* in, uint32_t* out, const Task* tasks) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t t = in[i];
  if (t == kNone) {
    out[i] = kNone;
    return;
  }
  const Task& tk = tasks[t];
  out[i] = i < tk.split ? tk.child_task[0] : tk.child_task[1];
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
      uint32_t id = atomicAdd(&ctr[0], 1u);
      uint32_t q = atomicAdd(&ctr[1], 1u);
      next[q] = Job{kid[s], id};
      o.child[s] = int32_t(id);
      o.count[s] = 0;
    }
  }
  out[jb.n4] = o;
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

unsigned Grid(uint64_t n, unsigned t = 256) { return unsigned((n + t - 1) / t); }

}  // namespace

void BuildBvh4Device(const float* bounds, const float4* records, uint32_t n, int maxLeaf, hipStream_t st,
                     DeviceBuffer& nodesOut, DeviceBuffer& primsOut, DeviceBvhStats& stats) {
  if (n == 0) throw std::runtime_error("scene has no primitives");
  if (maxLeaf < 1 || maxLeaf > 16) throw std::runtime_error("max_leaf_size must lie in [1, 16]");
  auto t0 = std::chrono::steady_clock::now();
  DeviceBuffer dB, dRec, dCb, dKeys, dKeys2, dIds, dIds2, dLo, dHi, dOwn0, dOwn1, dTasks0, dTasks1, dBins, dN2, dCtr, dTmp;
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
    dTmp.alloc(std::max<size_t>(tmpBytes, 16));
    HipCheck(hipcub::DeviceRadixSort::SortPairs(dTmp.ptr, tmpBytes, dKeys.as<uint64_t>(), dKeys2.as<uint64_t>(),
                                                dIds.as<uint32_t>(), dIds2.as<uint32_t>(), int(n), 0, 63, st),
             "radix sort");
  }
  const uint32_t* sorted = dIds2.as<uint32_t>();
  dLo.alloc(size_t(n) * 16);
  dHi.alloc(size_t(n) * 16);
  hipLaunchKernelGGL(k_gather_bounds, dim3(Grid(n)), dim3(256), 0, st, dB.as<float>(), sorted, n, dLo.as<float4>(),
                     dHi.as<float4>());
  // top-down levels
  const uint32_t maxNodes = 2 * n;
  dN2.alloc(size_t(maxNodes) * sizeof(N2));
  const uint32_t maxTasks = n / uint32_t(maxLeaf + 1) + 2;
  dTasks0.alloc(size_t(maxTasks) * sizeof(Task));
  dTasks1.alloc(size_t(maxTasks) * sizeof(Task));
  const uint32_t maxBins = n / kSmall + 2;
  dBins.alloc(size_t(maxBins) * kBins * sizeof(Bin));
  dOwn0.alloc(size_t(n) * 4);
  dOwn1.alloc(size_t(n) * 4);
  dCtr.alloc(64);
  BuildCtr* ctr = dCtr.as<BuildCtr>();
  uint32_t levels = 0;
  if (n <= uint32_t(maxLeaf)) {
    hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(64), 0, st, dLo.as<float4>(), dHi.as<float4>(), n, dN2.as<N2>());
  } else {
    Task root{0, 0, n, n > kSmall ? 0u : kNone, 0, {kNone, kNone}, 0};
    HipCheck(hipMemcpyAsync(dTasks0.ptr, &root, sizeof(Task), hipMemcpyHostToDevice, st), "root task");
    HipCheck(hipMemsetAsync(dOwn0.ptr, 0, size_t(n) * 4, st), "owners");
    BuildCtr c0{1, 0, 0, 0};
    HipCheck(hipMemcpyAsync(ctr, &c0, sizeof(c0), hipMemcpyHostToDevice, st), "counters");
    uint32_t nTasks = 1, nBins = n > kSmall ? 1 : 0;
    Task* cur = dTasks0.as<Task>();
    Task* nxt = dTasks1.as<Task>();
    uint32_t* own = dOwn0.as<uint32_t>();
    uint32_t* own2 = dOwn1.as<uint32_t>();
    while (nTasks > 0) {
      if (++levels > 128) throw std::runtime_error("BVH build did not converge");
      if (nBins) {
        hipLaunchKernelGGL(k_bins_clear, dim3(Grid(uint64_t(nBins) * kBins)), dim3(256), 0, st, dBins.as<Bin>(),
                           nBins * kBins);
        hipLaunchKernelGGL(k_bins_accum, dim3(Grid(n)), dim3(256), 0, st, n, own, cur, dLo.as<float4>(),
                           dHi.as<float4>(), dBins.as<Bin>());
      }
      HipCheck(hipMemsetAsync(&ctr->tasks, 0, 8, st), "level counters");
      hipLaunchKernelGGL(k_split, dim3(Grid(nTasks, 64)), dim3(64), 0, st, cur, nTasks, dBins.as<Bin>(), dLo.as<float4>(),
                         dHi.as<float4>(), dN2.as<N2>(), nxt, ctr, uint32_t(maxLeaf), Caps{maxNodes, maxTasks, maxBins});
      hipLaunchKernelGGL(k_owner, dim3(Grid(n)), dim3(256), 0, st, n, own, own2, cur);
      BuildCtr h{};
      HipCheck(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st), "level readback");
      HipCheck(hipStreamSynchronize(st), "BVH level");
      if (h.error || h.nodes > maxNodes || h.tasks > maxTasks) throw std::runtime_error("BVH build exceeded its buffers");
      nTasks = h.tasks;
      nBins = h.bins;
      std::swap(cur, nxt);
      std::swap(own, own2);
    }
  }
  // collapse
  BuildCtr hc{};
  HipCheck(hipMemcpyAsync(&hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, st), "node count");
  HipCheck(hipStreamSynchronize(st), "BVH build");
  const uint32_t n2count = n <= uint32_t(maxLeaf) ? 1u : hc.nodes;
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
  while (nJobs > 0) {
    depth++;
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
  primsOut.alloc(size_t(n) * 48);
  hipLaunchKernelGGL(k_gather_prims, dim3(Grid(n)), dim3(256), 0, st, dRec.as<float4>(), sorted, n, primsOut.as<float4>());
  // statistics: SAH cost relative to the root box
  BvhNode4 root;
  HipCheck(hipMemcpyAsync(&root, nodesOut.ptr, sizeof(root), hipMemcpyDeviceToHost, st), "root readback");
  HipCheck(hipStreamSynchronize(st), "BVH finish");
  float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
  for (int s = 0; s < 4; s++) {
    if (root.child[s] < 0) continue;
    lo[0] = std::min(lo[0], root.lox[s]), lo[1] = std::min(lo[1], root.loy[s]), lo[2] = std::min(lo[2], root.loz[s]);
    hi[0] = std::max(hi[0], root.hix[s]), hi[1] = std::max(hi[1], root.hiy[s]), hi[2] = std::max(hi[2], root.hiz[s]);
  }
  float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
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
  stats.levels = levels;
  stats.sah_cost = float(sah);
  stats.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace rad
