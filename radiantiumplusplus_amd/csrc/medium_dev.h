// Participating media on the device: the heterogeneous grid lookup
// (volume/grid.cpp:41-103), delta-tracking distance sampling
// (medium.cpp:19-43) and the phase functions (henyey_greenstein.cpp:18-40,
// isotropic.cpp:16-24).
#pragma once

#include "shading_dev.h"

namespace rad {

__device__ inline float wrap_coord(float v, int wrap) {
  if (wrap == RAD_WRAP_REPEAT) return fminf(fmaxf(v - floorf(v), 0.f), 1.f);
  return fminf(fmaxf(v, 0.f), 1.f);
}

__device__ inline f3 grid_read(const float* g, int sx, int sy, int ch, int u, int v, int w) {
  size_t start = ((size_t(w) * size_t(sy) + size_t(v)) * size_t(sx) + size_t(u)) * size_t(ch);
  if (ch == 1) return splat(g[start]);
  return mk(g[start], g[start + 1], g[start + 2]);
}

// The reference's "trilinear" lookup: scaled by (size - 1); the neighbour of
// each axis is chosen by the half-voxel offset and weighted by |frac - 0.5|.
__device__ inline f3 grid_eval(const float* g, int sx, int sy, int sz, int ch, int wrap, f3 p) {
  float wx = wrap_coord(p.x, wrap), wy = wrap_coord(p.y, wrap), wz = wrap_coord(p.z, wrap);
  float fx = wx * float(sx - 1), fy = wy * float(sy - 1), fz = wz * float(sz - 1);
  int ix = int(fx), iy = int(fy), iz = int(fz);
  int ax = min(max(ix + ((fx > (float(ix) + 0.5f)) ? 1 : -1), 0), sx - 1);
  int ay = min(max(iy + ((fy > (float(iy) + 0.5f)) ? 1 : -1), 0), sy - 1);
  int az = min(max(iz + ((fz > (float(iz) + 0.5f)) ? 1 : -1), 0), sz - 1);
  float du = fminf(fabsf(fx - float(ix) - 0.5f), 1.f);
  float dv = fminf(fabsf(fy - float(iy) - 0.5f), 1.f);
  float dw = fminf(fabsf(fz - float(iz) - 0.5f), 1.f);
  f3 c000 = grid_read(g, sx, sy, ch, ix, iy, iz), c100 = grid_read(g, sx, sy, ch, ax, iy, iz);
  f3 c010 = grid_read(g, sx, sy, ch, ix, ay, iz), c001 = grid_read(g, sx, sy, ch, ix, iy, az);
  f3 c110 = grid_read(g, sx, sy, ch, ax, ay, iz), c011 = grid_read(g, sx, sy, ch, ix, ay, az);
  f3 c101 = grid_read(g, sx, sy, ch, ax, iy, az), c111 = grid_read(g, sx, sy, ch, ax, ay, az);
  f3 x00 = fmadd(c000, splat(1.f - du), c100 * du);
  f3 x10 = fmadd(c010, splat(1.f - du), c110 * du);
  f3 y0 = fmadd(x00, splat(1.f - dv), x10 * dv);
  f3 x01 = fmadd(c001, splat(1.f - du), c101 * du);
  f3 x11 = fmadd(c011, splat(1.f - du), c111 * du);
  f3 y1 = fmadd(x01, splat(1.f - dv), x11 * dv);
  return fmadd(y0, splat(1.f - dw), y1 * dw);
}

}  // namespace rad
