// ORACLE / TEST INFRASTRUCTURE ONLY (see ref_math.h).
// extern "C" surface of the CPU restatement.  Consumed through ctypes by
// tests/ and by bench.py's cpu_baseline leg; never by the shipped library.
#pragma once
#include <stdint.h>

#include "../../include/rad_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ref_scene ref_scene;

typedef struct ref_render_params {
  int32_t integrator;  /* RAD_INTEGRATOR_PATH / RAD_INTEGRATOR_VOL_PATH */
  int32_t max_depth;
  int32_t rr_depth;
  uint32_t spp;
  uint32_t seed;
  int32_t threads;     /* 0 = all */
  int32_t tile_size;   /* 0 = 16 */
  int32_t x_begin, x_end, y_begin, y_end; /* sub-window (0 = full) */
  int32_t tile_step;   /* render tiles t = 0, step, 2*step, ... of the row-major tile list (0 = all) */
} ref_render_params;

const char* ref_last_error(void);
ref_scene* ref_scene_create(const rad_scene_desc* desc, int quirks);
void ref_scene_destroy(ref_scene* s);
double ref_scene_build_ms(const ref_scene* s);
void ref_scene_bounds(const ref_scene* s, float* lo, float* hi);
/* brute: 0 = the oracle's BVH, 1 = every primitive, 2 = every primitive with
   exact chunk culling (same result as 1) */
int ref_intersect_closest(const ref_scene* s, const rad_ray* rays, uint32_t n, int brute, rad_hit* hits);
int ref_intersect_any(const ref_scene* s, const rad_ray* rays, uint32_t n, int brute, uint8_t* occluded);
float ref_intersect_prim(const ref_scene* s, const rad_ray* ray, uint32_t geom, uint32_t prim);
void ref_camera_ray(const ref_scene* s, float sx, float sy, rad_ray* out);
int ref_render(const ref_scene* s, const ref_render_params* p, float* fb, float* sumsq, double* seconds);
// d[0] is probed; two_side / mask find their nested entry in d[0 .. count)
int ref_bsdf_eval(const rad_bsdf_desc* d, uint32_t count, int quirks, const float* in, uint32_t n, float* out);
int ref_bsdf_sample(const rad_bsdf_desc* d, uint32_t count, int quirks, const float* in, uint32_t n, float* out);
int ref_grid_eval(const rad_medium_desc* d, const float* in, uint32_t n, float* out);
float ref_erfinv(float a);
/* textures[index] of an array: in = n x {uv[2], duvdx[2], duvdy[2]}, out = n x 3 */
int ref_texture_eval(const rad_texture_desc* t, uint32_t count, uint32_t index, const float* in, uint32_t n, float* out);
/* first-hit uv and d(uv)/dx, d(uv)/dy of camera rays: in = n x {sx, sy}, out = n x 6 */
int ref_camera_footprint(const ref_scene* s, const float* in, uint32_t n, float* out);
float ref_medium_transmittance(const ref_scene* s, int medium, const rad_ray* ray, uint32_t channel, uint32_t seed);
/* phase functions (RAD_PHASE_*, g): eval in = n x {wi[3] (= -ray direction), wo[3]}, out = n;
   sample in = n x {wi[3], xi[2]}, out = n x {wo[3], pdf} */
int ref_phase_eval(int type, float g, const float* in, uint32_t n, float* out);
int ref_phase_sample(int type, float g, const float* in, uint32_t n, float* out);
/* Light::SampleDirection of scene light `light` from points: in = n x {p[3], xi[2]},
   out = n x {dir[3], dist, pdf, p_light[3], n_light[3], li[3], pdf_direction}, where the
   last is Light::PdfDirection evaluated on the returned sample */
int ref_light_sample(const ref_scene* s, uint32_t light, const float* in, uint32_t n, float* out);

#ifdef __cplusplus
}
#endif
