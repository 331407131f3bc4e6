// Host-owned state of a scene's guiding network: FP32 master weights, Adam
// moments, the FP16 / bias copies the kernels read, and the launchers.
#pragma once

#include <mutex>
#include <vector>

#include "guide.h"
#include "scene_host.h"

namespace rad {
namespace guide {

struct GuideState {
  DeviceBuffer params, grad, m1, m2, wh, bias, scal;
  // The FP16 weights and biases as of the last finished training pass (or
  // upload): renders that only read the guide copy these under pub_mu, so a
  // training render on another stream never hands them a half-updated net.
  DeviceBuffer pub_wh, pub_bias;
  std::mutex pub_mu;
  uint32_t steps = 0;
  uint64_t records_used = 0;
  GuideState();
  void Upload(const std::vector<float>& p);  // replace the weights, reset Adam
  void Publish(hipStream_t s);               // wh / bias -> pub_* once the work on s is done
  void Snapshot(void* wh_dst, void* bias_dst);  // the published copy -> device buffers of a render
  GuideNet Net(const float lo[3], const float hi[3], float alpha, float min_rough = 0.f) const;
  void Step(float lr, float gscale, hipStream_t s);  // Adam on the accumulated gradient
};

void InitParams(std::vector<float>& p, uint32_t seed);
void LaunchInfer(const GuideNet& net, const float* feats, const uint32_t* n_dev, uint32_t n_host, const uint32_t* slot_of,
                 _Float16* raw, int blocks, hipStream_t s);
void LaunchTrain(const GuideNet& net, const TrainRecord* recs, const uint32_t* batch, uint32_t B, float wnorm, float wmax,
                 float* grad, float* loss, hipStream_t s);
void LaunchBatch(uint32_t* batch, uint32_t B, uint32_t n, uint32_t seed, hipStream_t s);
void LaunchPdf(const float* raw, const float* dirs, uint32_t n, float* pdf, hipStream_t s);
void LaunchSample(const float* raw, const float* xi, uint32_t n, float* out, hipStream_t s);

}  // namespace guide
}  // namespace rad
