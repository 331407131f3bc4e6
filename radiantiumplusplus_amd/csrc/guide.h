// Learned path guiding: a 5-layer MLP (80 -> 64 -> 64 -> 64 -> 64 -> 48)
// maps a surface vertex (position, normal, outgoing direction; one-blob
// encoded) to an 8-lobe von Mises-Fisher mixture over world directions.
// Each lobe's mean direction points from the vertex to a target point the
// network outputs (parallax-aware lobes, as in Ruppert, Herholz and Lensch,
// "Robust fitting of parallax-aware mixtures for path guiding", 2020): light
// that arrives from a region of the scene, such as a lit spot, is one
// target for every vertex instead of a direction that turns with position.
// Inference and training run as FP16 MFMA kernels with FP32 accumulation;
// FP32 master weights are updated with Adam.  The mixture is combined with
// BSDF sampling by one-sample MIS, so guided renders stay unbiased.
//
// Parameter layout (FP32 master copy, one flat array): for each layer
// l = 0..4, W_l stored output-major ([out][in], row-major) followed by b_l.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace rad {
namespace guide {

constexpr int kLayers = 5;
constexpr int kIn = 80;    // 3 x 16 position bins + 2 x 8 normal bins + 2 x 8 direction bins
constexpr int kHid = 64;
constexpr int kOut = 48;   // 8 lobes x (logit, mean xyz, log kappa) = 40, padded to a 16 multiple
constexpr int kLobes = 8;
constexpr int kRawPerLobe = 5;
constexpr int kPosBins = 16, kDirBins = 8;
constexpr int kFeat = 9;   // raw per-vertex inputs: position, normal, outgoing direction

__host__ __device__ constexpr int LayerIn(int l) { return l == 0 ? kIn : kHid; }
__host__ __device__ constexpr int LayerOut(int l) { return l == kLayers - 1 ? kOut : kHid; }
__host__ __device__ constexpr int WeightOffset(int l) {  // offset of W_l in the flat FP32 array
  int o = 0;
  for (int i = 0; i < l; i++) o += LayerIn(i) * LayerOut(i) + LayerOut(i);
  return o;
}
__host__ __device__ constexpr int BiasOffset(int l) { return WeightOffset(l) + LayerIn(l) * LayerOut(l); }
constexpr int kParams = WeightOffset(kLayers);
__host__ __device__ constexpr int HalfOffset(int l) {  // offset of W_l in the FP16 weight array
  int o = 0;
  for (int i = 0; i < l; i++) o += LayerIn(i) * LayerOut(i);
  return o;
}
constexpr int kWeights = HalfOffset(kLayers);
constexpr int kBiases = 4 * kHid + kOut;

// Training record: raw features, sampled direction, the vertex throughput
// (before the bounce) and the radiance accumulated before the bounce's
// contribution; `w` is filled in when the path ends.
struct TrainRecord {
  float p[3], n[3], wi[3];
  float dir[3];
  float thr[3];
  float lbase[3];
  float w;        // F / p: luminance of (L_final - lbase) / thr
  uint32_t prev;  // previous record of the same path (0xffffffff = none)
};
static_assert(sizeof(TrainRecord) == 80, "record size");

// Device view of the network used by the inference kernel and the shading
// kernels.
struct GuideNet {
  const _Float16* w;  // kWeights halves, output-major per layer
  const float* b;     // kBiases
  float lo[3], inv_ext[3];  // scene bounds for the position encoding
  // Lobe targets: a lobe's three mean outputs t place a point c + h t
  // (c the bounds' centre, h their half extent per axis); the lobe points
  // from the vertex to it.
  float center[3], half[3];
  float alpha;        // probability of sampling the guide (vs the BSDF)
  float min_rough;    // glossy BSDFs with a smaller GGX/Beckmann alpha are not guided
  uint32_t max_depth; // vertices deeper than this many bounces are not guided
};

// Host-owned state of a scene's guide (weights, Adam moments, buffers).
struct GuideState;

}  // namespace guide
}  // namespace rad
