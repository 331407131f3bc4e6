// Wavefront state of a render job: structure-of-arrays path records in HBM,
// compacted work queues and per-slot accumulators.
//
// A job owns Q = pixels * K path slots; slot s always serves pixel s % pixels
// and its samples lane, lane + K, lane + 2K, ... (lane = s / pixels), so the
// per-slot accumulators can be reduced per pixel in a fixed order without
// atomics.  Each iteration runs
//   regen   (dead slots -> finalise the finished sample, start the next one)
//   extend  (closest hit for every queued path)
//   shade   (emission + MIS, NEE shadow-ray emission, BSDF sampling, RR)
//   shadow  (occlusion test, add unoccluded NEE contributions)
// with queue appends aggregated per wavefront (one atomic per 64 lanes).
#pragma once

#include <cstdint>

#include "scene_dev.h"

namespace rad {

enum : uint32_t { SAMPLE_NONE = 0xffffffffu, SAMPLE_DONE = 0xfffffffeu };
enum : uint32_t { PATH_SPECULAR = 1u };
enum : uint32_t { CNT_Q0 = 0, CNT_Q1 = 1, CNT_DEAD = 2, CNT_SHADOW = 3, CNT_WORDS = 8 };
enum : uint32_t { ST_SEGMENTS = 0, ST_SHADOW = 1, ST_SAMPLES = 2, ST_MEDIUM = 3, ST_WORDS = 8 };
enum : uint32_t { WF_QUIRKS = 1u };

struct Wavefront {
  float4* ray_o;   // origin, tmin
  float4* ray_d;   // direction, tmax
  float4* thr;     // throughput, eta
  float4* lsum;    // radiance of the sample in flight, previous BSDF pdf
  float4* prev;    // previous vertex position, path flags (bits)
  uint4* meta;     // pixel, sample index, depth, rng dimension
  uint32_t* key;   // rng key of the sample in flight
  uint32_t* medium;  // vol_path: current medium (+1, 0 = none) and channel in the high byte
  float4* acc;     // per-slot radiance sum
  float4* acc2;    // per-slot squared-radiance sum
  float4* hit;     // t, u, v, prim (bits)
  uint32_t* hit_geom;
  uint32_t* queue[2];
  uint32_t* dead;
  float4* sh_o;    // shadow origin, tmax
  float4* sh_d;    // shadow direction, slot (bits)
  float4* sh_c;    // shadow contribution
  uint32_t* cnt;   // CNT_* words
  unsigned long long* stats;  // ST_* words
  uint32_t npix, K, Q, spp;
  uint32_t seed, request;
  int32_t max_depth, rr_depth;
  uint32_t flags;
  int32_t integrator;
};

struct LaunchCfg {
  int blocks;
  int threads;
};

// Host-side launchers (defined in path_kernels.hip).
void LaunchInit(const Wavefront& w, hipStream_t s);
void LaunchRegen(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s);
void LaunchReset(const Wavefront& w, int out, hipStream_t s);
void LaunchExtend(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s);
void LaunchShadePath(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s);
void LaunchShadow(const Wavefront& w, const DevScene& scene, const LaunchCfg& c, hipStream_t s);
void LaunchResolve(const Wavefront& w, float* fb, float* fb_sumsq, hipStream_t s);

// Ray-batch entry points.
void LaunchClosest(const DevScene& scene, const void* rays, uint32_t n, void* hits, hipStream_t s);
void LaunchAny(const DevScene& scene, const void* rays, uint32_t n, uint8_t* occ, hipStream_t s);

// Per-function probes.
void LaunchBsdfEval(const BsdfRec& b, bool quirks, const float* in, uint32_t n, float* out, hipStream_t s);
void LaunchBsdfSample(const BsdfRec& b, bool quirks, const float* in, uint32_t n, float* out, hipStream_t s);

}  // namespace rad
