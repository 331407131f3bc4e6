// Device probes of the texture code the shading kernels run: texture
// lookups (Texture<T>::Eval) and the texture-space footprint of camera rays
// (SurfaceInteraction::BSDF(const RayDifferential&), interaction.cpp:83-111).
#include "texture_dev.h"
#include "wavefront.h"

namespace rad {

__global__ void k_tex_eval(DevScene s, uint32_t id, const float* in, uint32_t n, float* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* q = in + 6 * size_t(i);
  const TexCoord tc{q[0], q[1], q[2], q[3], q[4], q[5]};
  const f3 v = tex_eval(s, id, tc);
  out[3 * size_t(i)] = v.x;
  out[3 * size_t(i) + 1] = v.y;
  out[3 * size_t(i) + 2] = v.z;
}

// First hit of the camera ray through each screen position: uv, d(uv)/dx,
// d(uv)/dy (zeros where the ray misses).
__global__ void __launch_bounds__(kTraceBlock) k_footprint(DevScene s, const float* in, uint32_t n, float* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f3 o, d;
  float tmin, tmax;
  const float sx = in[2 * size_t(i)], sy = in[2 * size_t(i) + 1];
  camera_ray(s.cam, sx, sy, o, d, tmin, tmax);
  Hit h;
  const bool hit = trace<false>(s, o, d, tmin, tmax, h);
  float* r = out + 6 * size_t(i);
  TexCoord tc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (hit) {
    const SurfaceHit sh = surface_hit(s, o, d, h.t, h.u, h.v, h.prim, h.geom);
    tc = hit_texcoord(s, sh, make_float4(h.t, h.u, h.v, __uint_as_float(h.prim)), h.geom, true, sx, sy);
  }
  r[0] = tc.u, r[1] = tc.v, r[2] = tc.dudx, r[3] = tc.dvdx, r[4] = tc.dudy, r[5] = tc.dvdy;
}

void LaunchTexEval(const DevScene& s, uint32_t id, const float* in, uint32_t n, float* out, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_tex_eval, dim3((n + 127) / 128), dim3(128), 0, st, s, id, in, n, out);
}
void LaunchFootprint(const DevScene& s, const float* in, uint32_t n, float* out, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_footprint, dim3((n + kTraceBlock - 1) / kTraceBlock), dim3(kTraceBlock), 0, st, s, in, n, out);
}

}  // namespace rad
