// The guiding network's forward pass on FP16 MFMA: forward_tile for the
// training kernel (which keeps every layer's activations for the backward
// pass), infer_tile for the inference kernels (path vertices, probe inputs).
//
// Every layer is computed transposed, H_{l+1}^T = W_l H_l^T, with
// v_mfma_f32_16x16x16_f16 over 16 vertices per wave: the accumulator layout
// of a 16x16 output tile (lane = vertex + 16 * row group, 4 rows each) is
// exactly the B-operand layout of the next layer's K-chunk, so activations
// stay in registers between layers.
#pragma once

#include "guide_dev.h"

namespace rad {
namespace guide {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v mfma16(h4 a, h4 b, f4v c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }

__device__ __forceinline__ void load_weights_lds(_Float16* sW, float* sB, const _Float16* w, const float* b) {
  // kWeights halves as 8-byte chunks
  const uint2* src = reinterpret_cast<const uint2*>(w);
  uint2* dst = reinterpret_cast<uint2*>(sW);
  for (int i = threadIdx.x; i < kWeights / 4; i += blockDim.x) dst[i] = src[i];
  for (int i = threadIdx.x; i < kBiases; i += blockDim.x) sB[i] = b[i];
}

// Fragment-major copy for inference: the 16x16 A-operand tile (t, c) of
// layer l is 64 lanes x 4 halves, stored contiguously in lane order, so each
// MFMA operand is one conflict-free 8-byte LDS read per lane.  (Row-major
// rows of 64 or 80 halves put the 16 rows a wave reads on one or four
// banks.)
__device__ __forceinline__ int frag_base(int l, int t, int c) {
  return HalfOffset(l) + (t * (LayerIn(l) / 16) + c) * 256;
}

__device__ __forceinline__ void load_weights_lds_frag(_Float16* sW, _Float16* sBh, const _Float16* w, const float* b) {
  const uint2* src = reinterpret_cast<const uint2*>(w);
  for (int i = threadIdx.x; i < kWeights / 4; i += blockDim.x) {
    const int h = 4 * i;
    int l = 0;
#pragma unroll
    for (int k = 1; k < kLayers; k++) l += h >= HalfOffset(k) ? 1 : 0;
    const int nin = LayerIn(l);
    const int r = h - HalfOffset(l), o = r / nin, in = r - o * nin;
    const int lane = (o & 15) + 16 * ((in & 15) >> 2);
    *reinterpret_cast<uint2*>(sW + frag_base(l, o >> 4, in >> 4) + 4 * lane) = src[i];
  }
  for (int i = threadIdx.x; i < kBiases; i += blockDim.x) sBh[i] = (_Float16)b[i];
}

// A operand (16 rows 16t.., 16 columns 16c..) of layer l for this lane.
template <bool FRAG>
__device__ __forceinline__ h4 weight_operand(const _Float16* sW, int l, int t, int c, int lane) {
  if (FRAG) return *reinterpret_cast<const h4*>(sW + frag_base(l, t, c) + 4 * lane);
  const int v = lane & 15, g = lane >> 4;
  return *reinterpret_cast<const h4*>(sW + HalfOffset(l) + (16 * t + v) * LayerIn(l) + 16 * c + 4 * g);
}

__device__ __forceinline__ int bias_base(int l) { return l * kHid; }

// Forward pass of one 16-vertex tile for training.  fs: the lane's vertex
// scalars.  act[l][t]: post-ReLU outputs of hidden layer l (l = 0..3), tile
// t, in B-operand form; out[t]: the 48 raw outputs (3 tiles).  When
// xs != nullptr the layer-0 B operands (the one-blob inputs) are stored
// there as xs[chunk][lane].  FRAG: sW holds the fragment-major copy.
template <bool FRAG>
__device__ __forceinline__ void forward_tile(const _Float16* sW, const float* sB, const FeatScalars& fs, h4 act[4][4],
                                             f4v out[3], h4* xs) {
  const int lane = __lane_id();
  const int v = lane & 15, g = lane >> 4;
  f4v acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < kIn / 16; c++) {
    h4 b;
#pragma unroll
    for (int j = 0; j < 4; j++) b[j] = (_Float16)one_blob(fs, c, 4 * g + j);
    if (xs) xs[c * 64 + lane] = b;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      h4 a = weight_operand<FRAG>(sW, 0, t, c, lane);
      acc[t] = mfma16(a, b, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) act[0][t][j] = (_Float16)fmaxf(acc[t][j] + sB[bias_base(0) + 16 * t + 4 * g + j], 0.f);
#pragma unroll
  for (int l = 1; l < 4; l++) {
    __builtin_amdgcn_sched_barrier(0);  // keep each layer's weight loads inside the layer
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
      for (int t = 0; t < 4; t++) {
        h4 a = weight_operand<FRAG>(sW, l, t, c, lane);
        acc[t] = mfma16(a, act[l - 1][c], acc[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int j = 0; j < 4; j++) act[l][t][j] = (_Float16)fmaxf(acc[t][j] + sB[bias_base(l) + 16 * t + 4 * g + j], 0.f);
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int t = 0; t < 3; t++) out[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 4; c++) {
#pragma unroll
    for (int t = 0; t < 3; t++) {
      h4 a = weight_operand<FRAG>(sW, 4, t, c, lane);
      out[t] = mfma16(a, act[3][c], out[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < 3; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) out[t][j] += sB[bias_base(4) + 16 * t + 4 * g + j];
}

// Inference forward pass of one 16-vertex tile from the fragment-major
// weights (load_weights_lds_frag).  The epilogues run on packed FP16: an
// accumulator quad is converted to halves, then bias and ReLU are two packed
// adds and two packed maxes, a quarter of the per-value FP32 add / max /
// convert sequence (the layer outputs are FP16 MFMA operands either way).
// out[t]: raw outputs 16 t + 4 g .. 16 t + 4 g + 3 of vertex v.
__device__ __forceinline__ void infer_tile(const _Float16* sW, const _Float16* sBh, const FeatScalars& fs, h4 out[3]) {
  const int lane = __lane_id();
  const int g = lane >> 4;
  const h4 zero = h4{0, 0, 0, 0};
  f4v acc[4];
  h4 x[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < kIn / 16; c++) {
    h4 b;
#pragma unroll
    for (int j = 0; j < 4; j++) b[j] = (_Float16)one_blob(fs, c, 4 * g + j);
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = mfma16(weight_operand<true>(sW, 0, t, c, lane), b, acc[t]);
  }
#pragma unroll
  for (int l = 0; l < 4; l++) {
    if (l > 0) {
      __builtin_amdgcn_sched_barrier(0);  // keep each layer's weight loads inside the layer
#pragma unroll
      for (int t = 0; t < 4; t++) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 4; c++)
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = mfma16(weight_operand<true>(sW, l, t, c, lane), x[c], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const h4 bias = *reinterpret_cast<const h4*>(sBh + bias_base(l) + 16 * t + 4 * g);
      x[t] = __builtin_elementwise_max(__builtin_convertvector(acc[t], h4) + bias, zero);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  f4v o[3];
#pragma unroll
  for (int t = 0; t < 3; t++) o[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int t = 0; t < 3; t++) o[t] = mfma16(weight_operand<true>(sW, 4, t, c, lane), x[c], o[t]);
#pragma unroll
  for (int t = 0; t < 3; t++)
    out[t] = __builtin_convertvector(o[t], h4) + *reinterpret_cast<const h4*>(sBh + bias_base(4) + 16 * t + 4 * g);
}

}  // namespace guide
}  // namespace rad
