// FP16 MFMA kernels of the guiding network: inference for probe inputs, one
// training step (forward, loss, backward, weight gradients), and the Adam
// update of the FP32 master weights.  (Inference for path vertices is in
// the path kernels; both use the forward pass of guide_mlp.h.)
#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>
#include <vector>

#include "guide_dev.h"
#include "guide_mlp.h"
#include "guide_state.h"
#include "scene_host.h"

namespace rad {
namespace guide {

// Inference over n vertices: feats (9 floats per vertex) -> raw outputs
// (48 halves) at raw[slot_of[i] * 48] (or raw[i * 48] without slot_of).
__global__ void __launch_bounds__(256) k_guide_infer(GuideNet net, const float* feats, const uint32_t* n_dev, uint32_t n_host,
                                                     const uint32_t* slot_of, _Float16* raw) {
  __shared__ __align__(16) _Float16 sW[kWeights];
  __shared__ __align__(16) _Float16 sB[kBiases];
  load_weights_lds_frag(sW, sB, net.w, net.b);
  __syncthreads();
  const uint32_t n = n_dev ? *n_dev : n_host;
  const int lane = __lane_id();
  const int v = lane & 15, g = lane >> 4;
  const uint32_t wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  for (uint32_t tile = blockIdx.x * waves + wave; tile * 16 < n; tile += gridDim.x * waves) {
    uint32_t i = tile * 16 + v;
    float f9[kFeat];
    if (i < n) {
      for (int k = 0; k < kFeat; k++) f9[k] = feats[size_t(i) * kFeat + k];
    } else {
      for (int k = 0; k < kFeat; k++) f9[k] = 0.f;
    }
    FeatScalars fs = feat_scalars(f9, net.lo, net.inv_ext);
    h4 out[3];
    infer_tile(sW, sB, fs, out);
    if (i < n) {
      size_t base = size_t(slot_of ? slot_of[i] : i) * kOut;
#pragma unroll
      for (int t = 0; t < 3; t++) *reinterpret_cast<h4*>(raw + base + 16 * t + 4 * g) = out[t];
    }
  }
}

// Per-record loss gradient with respect to the 40 raw outputs, for
// loss_j = wn_j * (-log q(dir_j)), wn_j = min(w_j * wnorm, wmax).  The output
// gradients are kept in FP16 multiplied by kDzScale (undone in the weight
// gradient), so large sample weights stay inside the FP16 range.
constexpr float kDzScale = 1.f / 32.f;

// Lobe l points from the vertex p to the target c + h * (raw means).
__device__ inline float record_loss_grad(const float* raw, f3 dir, float wn, f3 p, f3 cen, f3 half, float* d) {
  float a[kLobes], logv[kLobes], kap[kLobes], c[kLobes], len[kLobes];
  f3 mu[kLobes];
  bool kin[kLobes];
  float mx = -kInf;
  for (int l = 0; l < kLobes; l++) mx = fmaxf(mx, raw[kRawPerLobe * l]);
  float se = 0.f;
  for (int l = 0; l < kLobes; l++) {
    a[l] = expf(raw[kRawPerLobe * l] - mx);
    se += a[l];
  }
  for (int l = 0; l < kLobes; l++) {
    a[l] /= se;  // pi_l
    f3 m = lobe_vector(raw + kRawPerLobe * l, p, cen, half);
    len[l] = sqrtf(sqnorm(m));
    mu[l] = len[l] > 1e-6f ? m / len[l] : mk(0.f, 1.f, 0.f);
    float kr = raw[kRawPerLobe * l + 4];
    kin[l] = kr > -3.f && kr < 6.f;
    kap[l] = expf(fminf(fmaxf(kr, -3.f), 6.f));
    c[l] = dot(mu[l], dir);
    logv[l] = logf(kap[l]) - logf(2.f * kPi) - log1pf(-expf(-2.f * kap[l])) + kap[l] * (c[l] - 1.f);
  }
  float lm = -kInf;
  for (int l = 0; l < kLobes; l++) lm = fmaxf(lm, logf(a[l]) + logv[l]);
  float s = 0.f;
  for (int l = 0; l < kLobes; l++) s += expf(logf(a[l]) + logv[l] - lm);
  float logq = lm + logf(s);
  for (int l = 0; l < kLobes; l++) {
    float r = expf(logf(a[l]) + logv[l] - logq);
    float k = kap[l];
    d[kRawPerLobe * l] = -wn * (r - a[l]);
    float e = expf(-2.f * k);
    float dlogv_dk = 1.f / k - 2.f * e / (1.f - e) + (c[l] - 1.f);
    d[kRawPerLobe * l + 4] = kin[l] ? -wn * r * dlogv_dk * k : 0.f;
    // d(log vMF) / d(target output) = (dir - mu c) kappa / |v| times h
    f3 gm = len[l] > 1e-6f ? mul(dir - mu[l] * c[l], half) * (k / len[l]) : splat(0.f);
    d[kRawPerLobe * l + 1] = -wn * r * gm.x;
    d[kRawPerLobe * l + 2] = -wn * r * gm.y;
    d[kRawPerLobe * l + 3] = -wn * r * gm.z;
  }
  return -wn * logq;
}

// One training step over 64 records per block (4 waves x 16).  Gradients
// (summed over records, not yet divided by the batch size) are added to
// grad[kParams]; the summed loss to *loss.
__global__ void __launch_bounds__(256) k_guide_train(GuideNet net, const TrainRecord* recs, const uint32_t* batch,
                                                     uint32_t B, float wnorm, float wmax, float* grad, float* loss) {
  __shared__ __align__(16) _Float16 sW[kWeights];
  __shared__ float sB[kBiases];
  __shared__ __align__(16) _Float16 sAct[kIn * 64 + 4 * kHid * 64];   // layer inputs A_0..A_4, [feature][record]
  __shared__ __align__(16) _Float16 sDz[4 * kHid * 64 + kOut * 64];   // output gradients dZ_0..dZ_4, [feature][record]
  __shared__ float sOut[64 * kOut];                     // raw outputs, [record][output]
  load_weights_lds(sW, sB, net.w, net.b);
  __syncthreads();
  const int lane = __lane_id();
  const int v = lane & 15, g = lane >> 4;
  const int wave = threadIdx.x >> 6;
  const int rloc = 16 * wave + v;  // record within the block
  const uint32_t ridx = blockIdx.x * 64 + rloc;
  const bool valid = ridx < B;
  TrainRecord rec{};
  if (valid) rec = recs[batch ? batch[ridx] : ridx];
  float f9[kFeat] = {rec.p[0], rec.p[1], rec.p[2], rec.n[0], rec.n[1], rec.n[2], rec.wi[0], rec.wi[1], rec.wi[2]};
  FeatScalars fs = feat_scalars(f9, net.lo, net.inv_ext);
  h4 act[4][4];
  f4v out[3];
  forward_tile<false>(sW, sB, fs, act, out, nullptr);
  auto actOff = [](int l) { return l == 0 ? 0 : kIn * 64 + (l - 1) * kHid * 64; };
  auto dzOff = [](int l) { return l * kHid * 64; };
  // layer inputs to LDS: X (recomputed one-blob values) and hidden activations
#pragma unroll
  for (int c = 0; c < kIn / 16; c++)
#pragma unroll
    for (int j = 0; j < 4; j++) sAct[actOff(0) + (16 * c + 4 * g + j) * 64 + rloc] = (_Float16)one_blob(fs, c, 4 * g + j);
#pragma unroll
  for (int l = 0; l < 4; l++)
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int j = 0; j < 4; j++) sAct[actOff(l + 1) + (16 * t + 4 * g + j) * 64 + rloc] = act[l][t][j];
#pragma unroll
  for (int t = 0; t < 3; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) sOut[rloc * kOut + 16 * t + 4 * g + j] = out[t][j];
  __syncthreads();
  // loss and output gradients: one lane per record
  if (g == 0) {
    float d[kOut];
    for (int k = 0; k < kOut; k++) d[k] = 0.f;
    float l = 0.f;
    if (valid && rec.w > 0.f) {
      float wn = fminf(rec.w * wnorm, wmax);
      l = record_loss_grad(sOut + rloc * kOut, mk(rec.dir[0], rec.dir[1], rec.dir[2]), wn,
                           mk(rec.p[0], rec.p[1], rec.p[2]), mk(net.center[0], net.center[1], net.center[2]),
                           mk(net.half[0], net.half[1], net.half[2]), d);
    }
    for (int k = 0; k < kOut; k++) sDz[dzOff(4) + k * 64 + rloc] = (_Float16)(d[k] * kDzScale);
    if (l != 0.f) atomicAdd(loss, l);
  }
  __syncthreads();
  // backward through the hidden layers of this wave's tile
  h4 dz[4];  // dZ of the layer above, in B-operand form
  {
    f4v acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      h4 b;
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = sDz[dzOff(4) + (16 * c + 4 * g + j) * 64 + rloc];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        h4 a;
#pragma unroll
        for (int j = 0; j < 4; j++) a[j] = sW[HalfOffset(4) + (16 * c + 4 * g + j) * kHid + 16 * t + v];
        acc[t] = mfma16(a, b, acc[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        _Float16 x = act[3][t][j] > (_Float16)0.f ? (_Float16)acc[t][j] : (_Float16)0.f;
        dz[t][j] = x;
        sDz[dzOff(3) + (16 * t + 4 * g + j) * 64 + rloc] = x;
      }
  }
#pragma unroll
  for (int l = 3; l >= 1; l--) {
    f4v acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
      for (int t = 0; t < 4; t++) {
        h4 a;
#pragma unroll
        for (int j = 0; j < 4; j++) a[j] = sW[HalfOffset(l) + (16 * c + 4 * g + j) * kHid + 16 * t + v];
        acc[t] = mfma16(a, dz[c], acc[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        _Float16 x = act[l - 1][t][j] > (_Float16)0.f ? (_Float16)acc[t][j] : (_Float16)0.f;
        dz[t][j] = x;
        sDz[dzOff(l - 1) + (16 * t + 4 * g + j) * 64 + rloc] = x;
      }
  }
  __syncthreads();
  // weight gradients: dW_l[out][in] = sum over the block's 64 records of dZ_l[out] A_l[in]
  int k = 0;
  for (int l = 0; l < kLayers; l++) {
    const int nin = LayerIn(l), nout = LayerOut(l);
    const int ti_n = nin / 16, to_n = nout / 16;
    for (int tile = 0; tile < ti_n * to_n; tile++, k++) {
      if ((k & 3) != wave) continue;
      const int to = tile / ti_n, ti = tile % ti_n;
      f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < 4; kc++) {
        h4 a = *reinterpret_cast<const h4*>(sDz + dzOff(l) + (16 * to + v) * 64 + 16 * kc + 4 * g);
        h4 b = *reinterpret_cast<const h4*>(sAct + actOff(l) + (16 * ti + v) * 64 + 16 * kc + 4 * g);
        acc = mfma16(a, b, acc);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        float x = acc[j];
        if (x != 0.f) atomicAdd(&grad[WeightOffset(l) + (16 * to + 4 * g + j) * nin + 16 * ti + v], x * (1.f / kDzScale));
      }
    }
    for (int o = threadIdx.x; o < nout; o += blockDim.x) {
      float sacc = 0.f;
      for (int r = 0; r < 64; r++) sacc += float(sDz[dzOff(l) + o * 64 + r]);
      if (sacc != 0.f) atomicAdd(&grad[BiasOffset(l) + o], sacc * (1.f / kDzScale));
    }
  }
}

// Adam on the FP32 master weights; refreshes the FP16 / bias copies the
// kernels read and clears the gradient.
__global__ void k_guide_adam(float* p, float* grad, float* m1, float* m2, _Float16* wh, float* bias, float lr, float b1t,
                             float b2t, float gscale) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
  float gi = grad[i] * gscale;
  float a = b1 * m1[i] + (1.f - b1) * gi;
  float c = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = c;
  float x = p[i] - lr * (a / (1.f - b1t)) / (sqrtf(c / (1.f - b2t)) + eps);
  p[i] = x;
  grad[i] = 0.f;
  for (int l = 0; l < kLayers; l++) {
    if (i >= WeightOffset(l) && i < BiasOffset(l)) wh[HalfOffset(l) + (i - WeightOffset(l))] = (_Float16)x;
    if (i >= BiasOffset(l) && i < WeightOffset(l + 1)) bias[l * kHid + (i - BiasOffset(l))] = x;
  }
}

// FP32 master -> FP16 / bias copies (no update).
__global__ void k_guide_publish(const float* p, _Float16* wh, float* bias) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  float x = p[i];
  for (int l = 0; l < kLayers; l++) {
    if (i >= WeightOffset(l) && i < BiasOffset(l)) wh[HalfOffset(l) + (i - WeightOffset(l))] = (_Float16)x;
    if (i >= BiasOffset(l) && i < WeightOffset(l + 1)) bias[l * kHid + (i - BiasOffset(l))] = x;
  }
}

// Mixture probes: pdf of directions, and samples, for raw output rows.
constexpr uint32_t kProbeBlock = 128;
__global__ void __launch_bounds__(kProbeBlock) k_guide_pdf(const float* raw, const float* dirs, uint32_t n, float* pdf) {
  __shared__ float4 km[kLobes * kProbeBlock];
  __shared__ float2 pw[kLobes * kProbeBlock];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Mixture m = lane_mixture(km, pw, threadIdx.x, kProbeBlock);
  decode(m, raw, splat(0.f), splat(0.f), splat(1.f));  // the means taken as directions
  pdf[i] = mixture_pdf(m, mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}
__global__ void __launch_bounds__(kProbeBlock) k_guide_sample(const float* raw, const float* xi, uint32_t n, float* out) {
  __shared__ float4 km[kLobes * kProbeBlock];
  __shared__ float2 pw[kLobes * kProbeBlock];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Mixture m = lane_mixture(km, pw, threadIdx.x, kProbeBlock);
  decode(m, raw, splat(0.f), splat(0.f), splat(1.f));
  f3 d = mixture_sample(m, xi[3 * i], xi[3 * i + 1], xi[3 * i + 2]);
  out[4 * i] = d.x;
  out[4 * i + 1] = d.y;
  out[4 * i + 2] = d.z;
  out[4 * i + 3] = mixture_pdf(m, d);
}

// Uniform random minibatch of record indices (with replacement).
__global__ void k_guide_batch(uint32_t* batch, uint32_t B, uint32_t n, uint32_t seed) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  batch[i] = mix32(mix32(seed) ^ (i * 0x9e3779b9u + 0x7f4a7c15u)) % n;
}

// ---------------------------------------------------------------- host state
void InitParams(std::vector<float>& p, uint32_t seed) {
  p.assign(kParams, 0.f);
  std::mt19937 rng(seed);
  for (int l = 0; l < kLayers; l++) {
    float lim = std::sqrt(6.f / float(LayerIn(l))) * (l == kLayers - 1 ? 0.1f : 1.f);
    std::uniform_real_distribution<float> U(-lim, lim);
    for (int i = 0; i < LayerIn(l) * LayerOut(l); i++) p[WeightOffset(l) + i] = U(rng);
  }
}

GuideState::GuideState() {
  params.alloc(kParams * 4);
  grad.alloc(kParams * 4);
  m1.alloc(kParams * 4);
  m2.alloc(kParams * 4);
  wh.alloc(kWeights * 2);
  bias.alloc(kBiases * 4);
  pub_wh.alloc(kWeights * 2);
  pub_bias.alloc(kBiases * 4);
  scal.alloc(64);
  std::vector<float> p;
  InitParams(p, 1234u);
  Upload(p);
}

void GuideState::Upload(const std::vector<float>& p) {
  if (p.size() != size_t(kParams)) throw std::runtime_error("guide parameter count mismatch");
  HipCheck(hipMemcpy(params.ptr, p.data(), kParams * 4, hipMemcpyHostToDevice), "guide params");
  HipCheck(hipMemset(grad.ptr, 0, kParams * 4), "guide grad");
  HipCheck(hipMemset(m1.ptr, 0, kParams * 4), "guide m1");
  HipCheck(hipMemset(m2.ptr, 0, kParams * 4), "guide m2");
  hipLaunchKernelGGL(k_guide_publish, dim3((kParams + 255) / 256), dim3(256), 0, nullptr, params.as<float>(), wh.as<_Float16>(),
                     bias.as<float>());
  HipCheck(hipGetLastError(), "guide publish");
  HipCheck(hipDeviceSynchronize(), "guide publish");
  Publish(nullptr);
  steps = 0;
}

void GuideState::Publish(hipStream_t s) {
  HipCheck(hipStreamSynchronize(s), "guide training");
  std::lock_guard<std::mutex> lk(pub_mu);
  HipCheck(hipMemcpyAsync(pub_wh.ptr, wh.ptr, kWeights * 2, hipMemcpyDeviceToDevice, s), "publish guide weights");
  HipCheck(hipMemcpyAsync(pub_bias.ptr, bias.ptr, kBiases * 4, hipMemcpyDeviceToDevice, s), "publish guide biases");
  HipCheck(hipStreamSynchronize(s), "publish guide");
}

void GuideState::Snapshot(void* wh_dst, void* bias_dst) {
  std::lock_guard<std::mutex> lk(pub_mu);
  HipCheck(hipMemcpy(wh_dst, pub_wh.ptr, kWeights * 2, hipMemcpyDeviceToDevice), "guide snapshot");
  HipCheck(hipMemcpy(bias_dst, pub_bias.ptr, kBiases * 4, hipMemcpyDeviceToDevice), "guide snapshot");
}

GuideNet GuideState::Net(const float lo[3], const float hi[3], float alpha, float min_rough) const {
  GuideNet n;
  n.w = wh.as<_Float16>();
  n.b = bias.as<float>();
  for (int i = 0; i < 3; i++) {
    float e = hi[i] - lo[i];
    n.lo[i] = lo[i];
    n.inv_ext[i] = e > 0.f ? 1.f / e : 0.f;
  }
  for (int i = 0; i < 3; i++) {
    n.center[i] = 0.5f * (lo[i] + hi[i]);
    n.half[i] = std::max(0.5f * (hi[i] - lo[i]), 1e-6f);
  }
  n.alpha = alpha;
  n.min_rough = min_rough;
  n.max_depth = 0xffffffffu;
  return n;
}

void LaunchInfer(const GuideNet& net, const float* feats, const uint32_t* n_dev, uint32_t n_host, const uint32_t* slot_of,
                 _Float16* raw, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_guide_infer, dim3(blocks), dim3(256), 0, s, net, feats, n_dev, n_host, slot_of, raw);
}

void LaunchTrain(const GuideNet& net, const TrainRecord* recs, const uint32_t* batch, uint32_t B, float wnorm, float wmax,
                 float* grad, float* loss, hipStream_t s) {
  if (B)
    hipLaunchKernelGGL(k_guide_train, dim3((B + 63) / 64), dim3(256), 0, s, net, recs, batch, B, wnorm, wmax, grad, loss);
}

void GuideState::Step(float lr, float gscale, hipStream_t s) {
  steps++;
  float b1t = std::pow(0.9f, float(steps)), b2t = std::pow(0.99f, float(steps));
  hipLaunchKernelGGL(k_guide_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, params.as<float>(), grad.as<float>(),
                     m1.as<float>(), m2.as<float>(), wh.as<_Float16>(), bias.as<float>(), lr, b1t, b2t, gscale);
}

void LaunchBatch(uint32_t* batch, uint32_t B, uint32_t n, uint32_t seed, hipStream_t s) {
  if (B && n) hipLaunchKernelGGL(k_guide_batch, dim3((B + 255) / 256), dim3(256), 0, s, batch, B, n, seed);
}

void LaunchPdf(const float* raw, const float* dirs, uint32_t n, float* pdf, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_guide_pdf, dim3((n + kProbeBlock - 1) / kProbeBlock), dim3(kProbeBlock), 0, s, raw, dirs, n, pdf);
}
void LaunchSample(const float* raw, const float* xi, uint32_t n, float* out, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_guide_sample, dim3((n + kProbeBlock - 1) / kProbeBlock), dim3(kProbeBlock), 0, s, raw, xi, n, out);
}

}  // namespace guide
}  // namespace rad

// ---------------------------------------------------------------- C ABI
namespace rad {
void SetLastError(const std::string& msg);
}

namespace {

using namespace rad;
using namespace rad::guide;

static_assert(sizeof(rad_guide_record) == sizeof(TrainRecord), "rad_guide_record must mirror TrainRecord");
static_assert(RAD_GUIDE_PARAMS == kParams && RAD_GUIDE_RAW == kOut && RAD_GUIDE_FEATS == kFeat, "guide ABI constants");

bool HaveDevice() {
  int count = 0;
  return hipGetDeviceCount(&count) == hipSuccess && count > 0;
}

template <class F>
rad_status Guarded(F&& f) {
  try {
    if (!HaveDevice()) {
      SetLastError("no HIP device available");
      return RAD_ERR_HIP;
    }
    f();
    return RAD_OK;
  } catch (const std::exception& e) {
    SetLastError(e.what());
    return RAD_ERR_HIP;
  }
}

}  // namespace

extern "C" {

rad_status rad_hip_guide_forward(const float* params, const float* bounds, const float* feats, uint32_t n, float* raw) {
  if (!params || !bounds || (n && (!feats || !raw))) {
    SetLastError("NULL argument");
    return RAD_ERR_INVALID_ARG;
  }
  return Guarded([&] {
    GuideState st;
    st.Upload(std::vector<float>(params, params + kParams));
    GuideNet net = st.Net(bounds, bounds + 3, 0.5f);
    DeviceBuffer df, dr;
    df.alloc(std::max<size_t>(16, size_t(n) * kFeat * 4));
    dr.alloc(std::max<size_t>(16, size_t(n) * kOut * 2));
    HipCheck(hipMemcpy(df.ptr, feats, size_t(n) * kFeat * 4, hipMemcpyHostToDevice), "copy features");
    if (n) LaunchInfer(net, df.as<float>(), nullptr, n, nullptr, dr.as<_Float16>(), std::min<uint32_t>(1024, (n + 63) / 64), nullptr);
    HipCheck(hipGetLastError(), "guide forward launch");
    std::vector<_Float16> h(size_t(n) * kOut);
    HipCheck(hipMemcpy(h.data(), dr.ptr, h.size() * 2, hipMemcpyDeviceToHost), "copy outputs");
    for (size_t i = 0; i < h.size(); i++) raw[i] = float(h[i]);
  });
}

rad_status rad_hip_guide_grad(const float* params, const float* bounds, const rad_guide_record* recs, uint32_t n,
                              float wnorm, float* grad, float* loss) {
  if (!params || !bounds || !grad || !loss || (n && !recs)) {
    SetLastError("NULL argument");
    return RAD_ERR_INVALID_ARG;
  }
  return Guarded([&] {
    GuideState st;
    st.Upload(std::vector<float>(params, params + kParams));
    GuideNet net = st.Net(bounds, bounds + 3, 0.5f);
    DeviceBuffer dr;
    dr.alloc(std::max<size_t>(16, size_t(n) * sizeof(TrainRecord)));
    HipCheck(hipMemcpy(dr.ptr, recs, size_t(n) * sizeof(TrainRecord), hipMemcpyHostToDevice), "copy records");
    HipCheck(hipMemset(st.scal.ptr, 0, 4), "clear loss");
    LaunchTrain(net, dr.as<TrainRecord>(), nullptr, n, wnorm, 64.f, st.grad.as<float>(), st.scal.as<float>(), nullptr);
    HipCheck(hipGetLastError(), "guide train launch");
    HipCheck(hipMemcpy(grad, st.grad.ptr, kParams * 4, hipMemcpyDeviceToHost), "copy gradient");
    float l = 0.f;
    HipCheck(hipMemcpy(&l, st.scal.ptr, 4, hipMemcpyDeviceToHost), "copy loss");
    float inv = n ? 1.f / float(n) : 0.f;
    for (int i = 0; i < kParams; i++) grad[i] *= inv;
    *loss = l * inv;
  });
}

rad_status rad_hip_guide_pdf(const float* raw, const float* dirs, uint32_t n, float* pdf) {
  if (!raw || (n && (!dirs || !pdf))) {
    SetLastError("NULL argument");
    return RAD_ERR_INVALID_ARG;
  }
  return Guarded([&] {
    DeviceBuffer dr, dd, dp;
    dr.alloc(kOut * 4);
    dd.alloc(std::max<size_t>(16, size_t(n) * 12));
    dp.alloc(std::max<size_t>(16, size_t(n) * 4));
    HipCheck(hipMemcpy(dr.ptr, raw, kRawPerLobe * kLobes * 4, hipMemcpyHostToDevice), "copy raw");
    HipCheck(hipMemcpy(dd.ptr, dirs, size_t(n) * 12, hipMemcpyHostToDevice), "copy dirs");
    LaunchPdf(dr.as<float>(), dd.as<float>(), n, dp.as<float>(), nullptr);
    HipCheck(hipGetLastError(), "guide pdf launch");
    HipCheck(hipMemcpy(pdf, dp.ptr, size_t(n) * 4, hipMemcpyDeviceToHost), "copy pdf");
  });
}

rad_status rad_hip_guide_sample(const float* raw, const float* xi, uint32_t n, float* out) {
  if (!raw || (n && (!xi || !out))) {
    SetLastError("NULL argument");
    return RAD_ERR_INVALID_ARG;
  }
  return Guarded([&] {
    DeviceBuffer dr, dx, dout;
    dr.alloc(kOut * 4);
    dx.alloc(std::max<size_t>(16, size_t(n) * 12));
    dout.alloc(std::max<size_t>(16, size_t(n) * 16));
    HipCheck(hipMemcpy(dr.ptr, raw, kRawPerLobe * kLobes * 4, hipMemcpyHostToDevice), "copy raw");
    HipCheck(hipMemcpy(dx.ptr, xi, size_t(n) * 12, hipMemcpyHostToDevice), "copy xi");
    LaunchSample(dr.as<float>(), dx.as<float>(), n, dout.as<float>(), nullptr);
    HipCheck(hipGetLastError(), "guide sample launch");
    HipCheck(hipMemcpy(out, dout.ptr, size_t(n) * 16, hipMemcpyDeviceToHost), "copy samples");
  });
}

rad_status rad_hip_scene_guide_get(rad_scene* scene, float* params, uint32_t* steps) {
  if (!scene || !params) {
    SetLastError("NULL argument");
    return RAD_ERR_INVALID_ARG;
  }
  return Guarded([&] {
    std::lock_guard<std::mutex> lk(scene->mu);
    HipCheck(hipSetDevice(scene->ctx->device), "hipSetDevice");
    if (!scene->guide) scene->guide = std::make_shared<GuideState>();
    HipCheck(hipMemcpy(params, scene->guide->params.ptr, kParams * 4, hipMemcpyDeviceToHost), "copy params");
    if (steps) *steps = scene->guide->steps;
  });
}

rad_status rad_hip_scene_guide_set(rad_scene* scene, const float* params) {
  if (!scene) {
    SetLastError("NULL argument");
    return RAD_ERR_INVALID_ARG;
  }
  return Guarded([&] {
    std::lock_guard<std::mutex> lk(scene->mu);
    HipCheck(hipSetDevice(scene->ctx->device), "hipSetDevice");
    if (!scene->guide) scene->guide = std::make_shared<GuideState>();
    std::vector<float> p;
    if (params)
      p.assign(params, params + kParams);
    else
      InitParams(p, 1234u);
    scene->guide->Upload(p);
  });
}

}  // extern "C"
