// Kernels of the participating-media path (renderer/vol_path.cpp) and the
// heterogeneous-grid probe.
#include <algorithm>
#include <stdexcept>

#include "medium_dev.h"
#include "scene_host.h"
#include "wavefront.h"

namespace rad {

__global__ void k_grid_eval(const float* grid, int sx, int sy, int sz, int ch, int wrap, const float* in, uint32_t n,
                            float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f3 v = grid_eval(grid, sx, sy, sz, ch, wrap, mk(in[3 * i], in[3 * i + 1], in[3 * i + 2]));
  out[3 * i] = v.x;
  out[3 * i + 1] = v.y;
  out[3 * i + 2] = v.z;
}

void SetLastError(const std::string& msg);

}  // namespace rad

extern "C" rad_status rad_hip_grid_eval(const rad_medium_desc* m, const float* in, uint32_t n, float* out) {
  using namespace rad;
  if (!m || !m->grid || (n && (!in || !out))) {
    SetLastError("NULL argument");
    return RAD_ERR_INVALID_ARG;
  }
  try {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
      SetLastError("no HIP device available");
      return RAD_ERR_HIP;
    }
    size_t gsz = size_t(m->grid_size[0]) * m->grid_size[1] * m->grid_size[2] * m->grid_channels;
    DeviceBuffer dg, din, dout;
    dg.alloc(std::max<size_t>(16, gsz * 4));
    din.alloc(std::max<size_t>(16, size_t(n) * 12));
    dout.alloc(std::max<size_t>(16, size_t(n) * 12));
    HipCheck(hipMemcpy(dg.ptr, m->grid, gsz * 4, hipMemcpyHostToDevice), "copy grid");
    HipCheck(hipMemcpy(din.ptr, in, size_t(n) * 12, hipMemcpyHostToDevice), "copy points");
    if (n)
      hipLaunchKernelGGL(k_grid_eval, dim3((n + 127) / 128), dim3(128), 0, nullptr, dg.as<float>(), m->grid_size[0],
                         m->grid_size[1], m->grid_size[2], m->grid_channels, m->grid_wrap, din.as<float>(), n,
                         dout.as<float>());
    HipCheck(hipGetLastError(), "grid probe launch");
    HipCheck(hipMemcpy(out, dout.ptr, size_t(n) * 12, hipMemcpyDeviceToHost), "copy results");
    return RAD_OK;
  } catch (const std::exception& e) {
    SetLastError(e.what());
    return RAD_ERR_HIP;
  }
}
