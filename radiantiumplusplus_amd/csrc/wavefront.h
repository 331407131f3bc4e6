// Wavefront state of a render job: structure-of-arrays path records in HBM,
// compacted work queues and the per-sample result entries.
//
// A job has room for Q paths in flight, not tied to pixels.  A path's state
// lives at its position in the extend queue it is in: queue q (the two
// alternate by iteration) holds positions [0, cnt[q]), and every per-path
// array holds 2 Q entries, state index q Q + position (rays and hits are
// per queue, by position).  Whoever appends a path to the next queue moves
// its state there, so every kernel reads and writes path state in queue
// order, never scattered by a slot number.  Camera samples are
// handed out from a global counter in sample-major order (id -> pixel
// id % pixels, sample row0 + id / pixels), so the queue stays full until
// the pass runs out of samples and the primary rays of a wave are adjacent
// pixels.  A finished sample's radiance goes to its own 16-byte result
// entry; after the pass, k_accum adds the entries of each pixel in sample
// order (the order of sample_renderer.cpp:97-105), so the image does not
// depend on which path ran which sample.  Each iteration runs
//   regen   (finished samples -> result entries, one new sample for each)
//   extend  (closest hit for every queued path)
//   shade   (emission + MIS, NEE shadow-ray emission, BSDF sampling, RR)
//   shadow  (occlusion test, add unoccluded NEE contributions)
// with queue appends aggregated per wavefront (one atomic per 64 lanes).
#pragma once

#include <cstdint>

#include "guide.h"
#include "scene_dev.h"
#include "traverse.h"

namespace rad {

enum : uint32_t { SAMPLE_NONE = 0xffffffffu };
enum : uint32_t { PATH_SPECULAR = 1u };
enum : uint32_t {
  CNT_Q0 = 0,
  CNT_Q1 = 1,
  CNT_DEAD = 2,
  CNT_SHADOW = 3,
  CNT_NEXT = 4,       // next sample id of the pass
  CNT_FETCH_EXT = 5,  // extend-kernel work counter
  CNT_FETCH_SH = 6,   // shadow-kernel work counter
  CNT_MAT = 8,        // 5 words: per-material queue counts (diffuse + misses, metal, glass, Disney, others)
  CNT_SUSP = 14,      // 2 words: traversals suspended into resume pool 0 / 1
  CNT_GMAT = 16,      // 5 words: guided vertices per material (stored from the back of that material's queue)
  CNT_WORDS = 24      // even, so the 64-bit statistics that follow stay aligned
};
enum : uint32_t { ST_SAMPLES = 0, ST_MEDIUM = 1, ST_QCHECK = 2, ST_QERR = 3, ST_WORDS = 4 };
enum : uint32_t { WF_QUIRKS = 1u, WF_GUIDE = 2u, WF_RECORD = 4u };
enum : uint32_t { G_QUEUE = 0, G_RECORDS = 1, G_WSUM = 2, G_WORDS = 4 };
enum : uint32_t { MED_UNSET = 0xffffffffu };

// Sorted shading queues: BSDF types 0..3 (diffuse, rough metal, rough glass,
// Disney) have their own; kMatOther takes every other type.
constexpr int kMatOther = 4;
constexpr uint32_t kNoSusp = 0xffffffffu;
// Dead-list flag: the shading kernel already wrote this path's finished
// sample to the result entries, so regen reads nothing for it.
constexpr uint32_t kDeadDone = 0x80000000u;
constexpr int kMatQueues = 5;

// State index of position i of extend queue q.
__device__ __forceinline__ uint32_t state_index(uint32_t Q, int q, uint32_t i) { return uint32_t(q) * Q + i; }

// Row of entry j of material k's guided queue in Wavefront::graw: the
// materials' guided vertices one after another, in material order.
__device__ __forceinline__ uint32_t guided_row(const uint32_t* cnt, int k, uint32_t j) {
  uint32_t r = j;
  for (int q = 0; q < k; q++) r += cnt[CNT_GMAT + q];
  return r;
}

// Field k of a 64-byte per-path record of four 16-byte fields: element s is
// p[4 s], so the four fields of a path share one memory sector.
template <class T>
struct SlotField {
  T* p;
  __host__ __device__ T& operator[](size_t s) const { return p[4 * s]; }
};

struct Wavefront {
  // Ray at position i of extend queue q: ray[q][2 i] = origin, tmin;
  // ray[q][2 i + 1] = direction, tmax (one 32-byte record, so a ray is one
  // memory sector, not two).  Stored by queue position, not by slot: regen
  // and shading write the rays they append in block-contiguous runs, and
  // extend, shading and inference read them in queue order.
  float4* ray[2];
  // the path record at state index s: four fields in one 64-byte record
  SlotField<float4> thr;   // throughput, eta
  SlotField<float4> lsum;  // radiance of the sample in flight, previous BSDF pdf
  SlotField<float4> prev;  // previous vertex position, path flags (bits)
  SlotField<uint4> meta;   // pixel, sample index (SAMPLE_NONE = idle), depth, rng dimension
  uint32_t* key;   // rng key of the sample in flight (by state index)
  uint32_t* medium;  // vol_path, by state index: MED_UNSET, or current medium + 1 (0 = none) | channel << 24
  // hit of the ray at position i of the extend queue (not slot-indexed, so
  // extend writes and classify reads consecutive entries)
  float4* hit;     // t, u, v, prim (bits)
  uint32_t* hit_geom;
  // Traversals cut off at visit_budget node visits (path jobs; null = never):
  // by state index the resume-pool index, or kNoSusp.  Iteration `in` resumes from
  // resume[in] and suspends into resume[1 - in] (count in CNT_SUSP + 1 - in);
  // a suspended path skips shading and goes straight to the next extend queue.
  uint32_t* susp;
  ResumeState* resume[2];
  uint32_t resume_cap, visit_budget;
  // The extend queue partitioned by material (k_classify), as extend-queue
  // positions; [4]: every other type.  With guiding or training records on, entry j of a material's
  // guided vertices is mq[k][Q - 1 - j] (count CNT_GMAT + k): both parts
  // together never exceed the Q paths.
  uint32_t* mq[5];
  uint32_t* dead;  // state indices of the samples that ended this iteration (| kDeadDone: result already written)
  uint32_t* passsrc;  // queue check only: extend-queue position each passed-through entry came from
  float4* sh_o;    // shadow origin, tmax
  float4* sh_d;    // shadow direction, state index of the path it adds to (bits)
  float4* sh_c;    // shadow contribution
  float4* res;     // result entries: RGB (+ padding) of rows * npix samples
  float4* pacc;    // per-pixel radiance sum
  float4* pacc2;   // per-pixel squared-radiance sum
  uint32_t* cnt;   // CNT_* words
  unsigned long long* stats;  // ST_* words
  // learned guiding (null when off)
  uint32_t* gq;      // unsorted shading: extend-queue positions of this iteration's guided vertices
  // 48 raw network outputs (40 used) per guided vertex.  Sorted shading: row
  // guided_row(cnt, k, j) for entry j of material k's guided queue, so
  // inference writes and shading reads consecutive rows; unsorted: row =
  // extend-queue position.
  _Float16* graw;
  uint32_t* gpend;   // by state index: training record still waiting for its radiance baseline
  uint32_t* glast;   // by state index: newest training record of the path in flight
  guide::TrainRecord* recs;
  uint32_t* gcnt;    // G_* words (G_WSUM holds float bits)
  uint32_t rec_cap, rec_stride;
  guide::GuideNet net;
  uint32_t npix, Q, spp;
  uint32_t row0, rows;   // samples [row0, row0 + rows) of every pixel form the current pass
  uint32_t seed, request;
  int32_t max_depth, rr_depth;
  uint32_t flags;
  int32_t integrator;
};

// Append `pred` lanes of this wavefront to a queue with one atomic per wave.
// Must be reached by every lane of the wave (uniform control flow).
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return 0;
  uint32_t lane = __lane_id();
  uint32_t leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ void wave_count(unsigned long long* counter, bool pred) {
  unsigned long long mask = __ballot(pred);
  if (mask != 0ull && __lane_id() == (uint32_t)(__ffsll((long long)mask) - 1))
    atomicAdd(counter, (unsigned long long)__popcll(mask));
}

// Block-wide queue appends: thread t appends when pred[k], to counter k, for
// up to N counters at once; one atomic per counter per block (the waves'
// counts are combined in LDS).  Every thread of the block must call it with
// the same N (it synchronises the block); lds holds at least N * (waves + 1)
// words.
template <int N>
__device__ __forceinline__ void block_append(uint32_t* const (&ctr)[N], const bool (&pred)[N], uint32_t (&pos)[N],
                                             uint32_t* lds) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long mask[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    mask[k] = __ballot(pred[k]);
    if (lane == 0) lds[k * W + wave] = (uint32_t)__popcll(mask[k]);
  }
  __syncthreads();
  if (threadIdx.x < uint32_t(N)) {
    const int k = int(threadIdx.x);
    uint32_t total = 0;
    for (uint32_t v = 0; v < W; v++) {
      uint32_t c = lds[k * W + v];
      lds[k * W + v] = total;
      total += c;
    }
    lds[N * W + k] = total ? atomicAdd(ctr[k], total) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) pos[k] = lds[N * W + k] + lds[k * W + wave] + (uint32_t)__popcll(mask[k] & below);
  __syncthreads();
}

// Block-wide append of E entries per thread.  Entry e goes to one of the
// first N - EXTRA counters (cls[e], or -1 for none) and, with EXTRA, also to
// counter N - 1 when ext[e]; one atomic per counter covers all E entries of
// every thread, so the block's load -> count -> atomic -> store chain is paid
// once per E * blockDim entries.  pos / pos_ext: each entry's position in its
// queue(s).  lds: N * (waves + 1) words.  Every thread of the block calls it.
template <int N, int E, bool EXTRA>
__device__ __forceinline__ void block_append_n(uint32_t* const (&ctr)[N], const int (&cls)[E], const bool (&ext)[E],
                                               uint32_t (&pos)[E], uint32_t (&pos_ext)[E], uint32_t* lds) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int k = 0; k < N; k++) {
    const bool isExt = EXTRA && k == N - 1;
    uint32_t acc = 0;  // wave-uniform running count of counter k
#pragma unroll
    for (int e = 0; e < E; e++) {
      const bool p = isExt ? ext[e] : cls[e] == k;
      const unsigned long long m = __ballot(p);
      const uint32_t local = acc + (uint32_t)__popcll(m & below);
      if (p) {
        if (isExt)
          pos_ext[e] = local;
        else
          pos[e] = local;
      }
      acc += (uint32_t)__popcll(m);
    }
    if (lane == 0) lds[k * W + wave] = acc;
  }
  __syncthreads();
  if (threadIdx.x < uint32_t(N)) {
    const int k = int(threadIdx.x);
    uint32_t total = 0;
    for (uint32_t v = 0; v < W; v++) {
      uint32_t c = lds[k * W + v];
      lds[k * W + v] = total;
      total += c;
    }
    lds[N * W + k] = total ? atomicAdd(ctr[k], total) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; e++) {
    if (cls[e] >= 0) pos[e] += lds[N * W + cls[e]] + lds[cls[e] * W + wave];
    if (EXTRA && ext[e]) pos_ext[e] += lds[N * W + N - 1] + lds[(N - 1) * W + wave];
  }
  __syncthreads();
}

// Block-wide sum into a 64-bit statistics counter (one atomic per block).
__device__ __forceinline__ void block_count(unsigned long long* counter, uint32_t v, uint32_t* lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const uint32_t wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  if (__lane_id() == 0) lds[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t k = 0; k < W; k++) t += lds[k];
    if (t) atomicAdd(counter, t);
  }
  __syncthreads();
}

struct LaunchCfg {
  int blocks;        // persistent grid of the per-path kernels (regen, shade, classify)
  int trace_blocks;  // persistent grid of the traversal kernels (extend, shadow)
};

// Host-side launchers (defined in path_kernels.hip).
void LaunchInitPass(const Wavefront& w, bool first, hipStream_t s);
void LaunchRegen(const Wavefront& w, const DevScene& scene, int in, bool fresh, const LaunchCfg& c, hipStream_t s);
void LaunchReset(const Wavefront& w, int out, hipStream_t s);
void LaunchExtend(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s);
void LaunchShadePath(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s);
// Material-sorted shading: classify buckets the extend queue by BSDF type
// (with WF_GUIDE it also fills the guide queue), then one specialised shade
// kernel per type present (mat_mask bit = BSDF type).
void LaunchClassify(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s);
// Queue check after a classify pass (tally: Q zeroed words, left zeroed).
void LaunchQueueCheck(const Wavefront& w, const DevScene& scene, int in, uint32_t* tally, bool selftest,
                      const LaunchCfg& c, hipStream_t s);
void LaunchShadeMaterials(const Wavefront& w, const DevScene& scene, int in, uint32_t mat_mask, const LaunchCfg& c,
                          hipStream_t s);
void LaunchShadow(const Wavefront& w, const DevScene& scene, const LaunchCfg& c, hipStream_t s);
void LaunchGuidePrep(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s);
void LaunchGuideInfer(const Wavefront& w, const DevScene& scene, int in, int blocks, bool sorted, hipStream_t s);
void LaunchAccum(const Wavefront& w, hipStream_t s);
// vol_path (defined in volpath_kernels.hip).
void LaunchShadeVol(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s);
void LaunchShadeVolMaterials(const Wavefront& w, const DevScene& scene, int in, uint32_t mask, const LaunchCfg& c,
                             hipStream_t s);
void LaunchVolShadow(const Wavefront& w, const DevScene& scene, const LaunchCfg& c, hipStream_t s);
void LaunchFinal(const Wavefront& w, float* fb, float* fb_sumsq, hipStream_t s);

// Ray-batch entry points.
void LaunchClosest(const DevScene& scene, const void* rays, uint32_t n, void* hits, hipStream_t s);
void LaunchAny(const DevScene& scene, const void* rays, uint32_t n, uint8_t* occ, hipStream_t s);

// Per-function probes.
void LaunchTexEval(const DevScene& s, uint32_t id, const float* in, uint32_t n, float* out, hipStream_t st);
void LaunchFootprint(const DevScene& s, const float* in, uint32_t n, float* out, hipStream_t st);
void LaunchBsdfEval(const BsdfRec& b, bool quirks, const float* in, uint32_t n, float* out, hipStream_t s);
void LaunchBsdfSample(const BsdfRec& b, bool quirks, const float* in, uint32_t n, float* out, hipStream_t s);

}  // namespace rad
