// Wavefront kernels of the `path` integrator (renderer/path.cpp:131-227) and
// the ray-batch intersection kernels (accel/embree.cpp:79-111).
#include <algorithm>

#include "shading_dev.h"
#include "wavefront.h"

namespace rad {

// Start of a pass: every slot idle and on the dead list, so the first regen
// hands each one a sample.
__global__ void k_init_pass(Wavefront w, int first) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (first && s < w.npix) {
    w.pacc[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    w.pacc2[s] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (s < w.Q) {
    w.meta[s] = make_uint4(0u, SAMPLE_NONE, 0u, 0u);
    w.dead[s] = s;
  }
  if (s == 0) {
    w.cnt[CNT_Q0] = 0;
    w.cnt[CNT_Q1] = 0;
    w.cnt[CNT_DEAD] = w.Q;
    w.cnt[CNT_SHADOW] = 0;
    w.cnt[CNT_NEXT] = 0;
    w.cnt[CNT_FETCH_EXT] = 0;
    w.cnt[CNT_FETCH_SH] = 0;
    if (first)
      for (int i = 0; i < ST_WORDS; i++) w.stats[i] = 0ull;
  }
}

// Store the sample a dead slot finished into the result planes and start the
// next camera sample of the pass on it (sample_renderer.cpp:97-99).
__global__ void __launch_bounds__(256) k_regen(Wavefront w, DevScene scene, int in) {
  const uint32_t n = w.cnt[CNT_DEAD];
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t pass = w.rows * w.npix;
  const size_t plane = size_t(pass);
  const bool more = w.cnt[CNT_NEXT] < pass;  // stale is fine: only skips atomics once exhausted
  const uint32_t W = uint32_t(scene.cam.width);
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < n;
    bool finished = false;
    uint32_t s = 0;
    if (valid) {
      s = w.dead[i];
      uint4 m = w.meta[s];
      if (m.y != SAMPLE_NONE) {
        float4 l = w.lsum[s];
        size_t r = size_t(m.y - w.row0) * w.npix + m.x;
        w.res[r] = l.x;
        w.res[r + plane] = l.y;
        w.res[r + 2 * plane] = l.z;
        finished = true;
      }
    }
    uint32_t id = more ? wave_append(&w.cnt[CNT_NEXT], valid) : pass;
    bool start = valid && id < pass;
    if (start) {
      uint32_t pixel = id % w.npix, sample = w.row0 + id / w.npix;
      uint32_t key = sample_key(w.seed, w.request, pixel, sample);
      uint32_t px = pixel % W, py = pixel / W;
      float jx = rng_u(key, 0), jy = rng_u(key, 1);
      f3 o, d;
      float tmin, tmax;
      camera_ray(scene.cam, float(px) + jx, float(py) + jy, o, d, tmin, tmax);
      w.ray_o[s] = make_float4(o.x, o.y, o.z, tmin);
      w.ray_d[s] = make_float4(d.x, d.y, d.z, tmax);
      w.thr[s] = make_float4(1.f, 1.f, 1.f, 1.f);
      w.lsum[s] = make_float4(0.f, 0.f, 0.f, 1.f);
      w.prev[s] = make_float4(0.f, 0.f, 0.f, __uint_as_float(PATH_SPECULAR));
      w.meta[s] = make_uint4(pixel, sample, 0u, 2u);
      w.key[s] = key;
      if (w.medium) w.medium[s] = 0xffffffffu;  // vol_path: medium / channel chosen at shade time
    } else if (valid) {
      w.meta[s] = make_uint4(0u, SAMPLE_NONE, 0u, 0u);  // idle for the rest of the pass
    }
    uint32_t pos = wave_append(&w.cnt[in], start);
    if (start) w.queue[in][pos] = s;
    wave_count(&w.stats[ST_SAMPLES], finished);
  }
}

__global__ void k_reset(uint32_t* cnt, int out) {
  if (threadIdx.x == 0) {
    cnt[out] = 0;
    cnt[CNT_DEAD] = 0;
    cnt[CNT_SHADOW] = 0;
    cnt[CNT_FETCH_EXT] = 0;
    cnt[CNT_FETCH_SH] = 0;
  }
}

struct ExtendJob {
  const uint32_t* queue;
  const float4 *ray_o, *ray_d;
  float4* hit;
  uint32_t* hit_geom;
  __device__ __forceinline__ bool load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax, uint32_t& tag) {
    uint32_t s = queue[i];
    float4 a = ray_o[s], b = ray_d[s];
    o = mk(a.x, a.y, a.z);
    d = mk(b.x, b.y, b.z);
    tmin = a.w;
    tmax = b.w;
    tag = s;
    return true;
  }
  __device__ __forceinline__ void finish(uint32_t, uint32_t s, const TraceRay& r) {
    hit[s] = make_float4(r.h.t, r.h.u, r.h.v, __uint_as_float(r.hit ? r.h.prim : 0u));
    hit_geom[s] = r.hit ? r.h.geom : 0xffffffffu;
  }
};

__global__ void __launch_bounds__(kTraceBlock) k_extend(Wavefront w, DevScene scene, int in) {
  ExtendJob job{w.queue[in], w.ray_o, w.ray_d, w.hit, w.hit_geom};
  trace_stream<false>(scene, w.cnt[in], &w.cnt[CNT_FETCH_EXT], job);
}

// One bounce of Path::Li (path.cpp:141-225) for every queued path.
__global__ void __launch_bounds__(256) k_shade_path(Wavefront w, DevScene scene, int in) {
  const uint32_t n = w.cnt[in];
  const uint32_t stride = gridDim.x * blockDim.x;
  const bool quirks = (w.flags & WF_QUIRKS) != 0;
  const int out = 1 - in;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    bool alive = false, emit = false;
    uint32_t s = 0;
    float4 so = make_float4(0, 0, 0, 0), sdir = make_float4(0, 0, 0, 0), sc = make_float4(0, 0, 0, 0);
    if (i < n) {
      s = w.queue[in][i];
      float4 ro4 = w.ray_o[s], rd4 = w.ray_d[s];
      f3 ro = mk(ro4.x, ro4.y, ro4.z), rd = mk(rd4.x, rd4.y, rd4.z);
      float4 hh = w.hit[s];
      uint32_t geom = w.hit_geom[s];
      float4 t4 = w.thr[s];
      f3 T = mk(t4.x, t4.y, t4.z);
      float eta = t4.w;
      float4 l4 = w.lsum[s];
      f3 L = mk(l4.x, l4.y, l4.z);
      float prevPdf = l4.w;
      float4 p4 = w.prev[s];
      f3 prevP = mk(p4.x, p4.y, p4.z);
      uint32_t flags = __float_as_uint(p4.w);
      bool spec = flags & PATH_SPECULAR;
      uint4 m = w.meta[s];
      uint32_t depth = m.z, dim = m.w;
      uint32_t key = w.key[s];
      if (geom == 0xffffffffu) {
        // miss: environment emission (scene.cpp:105-113, skybox.cpp:53-63)
        if (scene.env_light >= 0) {
          float wgt = 1.f;
          if (!spec) wgt = mis_weight(prevPdf, light_pdf_dir(scene, scene.env_light, prevP, rd, rd, rd, 0.f));
          L += mul(T, scene.lights[scene.env_light].radiance) * wgt;
        }
      } else {
        SurfaceHit si = surface_hit(scene, ro, rd, hh.x, hh.y, hh.z, __float_as_uint(hh.w), geom);
        const ShapeRec& sr = scene.shapes[geom];
        if (sr.light >= 0) {  // emitter hit with MIS (path.cpp:146-164)
          float wgt = 1.f;
          if (!spec) {
            f3 rel = si.p - prevP;
            float dist = norm(rel);
            f3 dir = rel / dist;
            wgt = mis_weight(prevPdf, light_pdf_dir(scene, sr.light, prevP, si.p, si.sh.n, dir, dist));
          }
          f3 le = dot(si.sh.to_world(si.wi), si.n) >= 0.f ? scene.lights[sr.light].radiance : splat(0.f);
          L += mul(T, le) * wgt;
        }
        bool stop = (w.max_depth >= 0 && int32_t(depth) + 1 >= w.max_depth) || sr.light >= 0;
        if (!stop) {
          const BsdfRec& b = scene.bsdfs[sr.bsdf];
          if (b.lobes & LOBE_NO_DELTA) {  // NEE (path.cpp:173-191)
            float ul = rng_u(key, dim), u1 = rng_u(key, dim + 1), u2 = rng_u(key, dim + 2);
            dim += 3;
            LightSample ls = sample_light(scene, si.p, ul, u1, u2);
            if (ls.pdf > 0.f) {
              f3 wo = si.sh.to_local(ls.dir);
              float bp;
              f3 f = bsdf_eval(b, quirks, si.wi, wo, bp);
              // Scene::IsOcclude setup (scene.cpp:53-63)
              f3 o = offset_p(si.p, si.n, ls.p - si.p);
              f3 dd = ls.p - o;
              float dist = norm(dd);
              if (!(dist < kShadowEps)) {
                dd = dd / dist;
                float wgt = ls.delta ? 1.f : mis_weight(ls.pdf, bp);
                f3 c = mul(mul(T, f), ls.li) * wgt / ls.pdf;
                if (!is_zero(c)) {
                  emit = true;
                  so = make_float4(o.x, o.y, o.z, dist * (1.f - kShadowEps));
                  sdir = make_float4(dd.x, dd.y, dd.z, __uint_as_float(s));
                  sc = make_float4(c.x, c.y, c.z, 0.f);
                }
              }
            }
          }
          // BSDF sampling (path.cpp:192-210)
          float ul = rng_u(key, dim), u1 = rng_u(key, dim + 1), u2 = rng_u(key, dim + 2);
          dim += 3;
          BsdfSample bs;
          f3 f = bsdf_sample(b, quirks, si.wi, ul, u1, u2, bs);
          if (bs.pdf > 0.f) {
            T = mul(T, f) / bs.pdf;
            eta *= bs.eta;
            f3 wo = si.sh.to_world(bs.wo);
            f3 no = offset_p(si.p, si.n, wo);
            w.ray_o[s] = make_float4(no.x, no.y, no.z, 0.f);
            w.ray_d[s] = make_float4(wo.x, wo.y, wo.z, kFloatMax);
            prevPdf = bs.pdf;
            prevP = si.p;
            spec = (bs.lobes & LOBE_DELTA) != 0;
            alive = true;
            if (int32_t(depth) >= w.rr_depth) {  // Russian roulette (path.cpp:212-224)
              float rr = fminf(maxc(T) * sqr(eta), 0.95f);
              if (rng_u(key, dim) > rr) {
                alive = false;
              } else {
                T = T * (1.f / rr);
              }
              dim += 1;
            }
          }
        }
      }
      w.thr[s] = make_float4(T.x, T.y, T.z, eta);
      w.lsum[s] = make_float4(L.x, L.y, L.z, prevPdf);
      w.prev[s] = make_float4(prevP.x, prevP.y, prevP.z, __uint_as_float(spec ? PATH_SPECULAR : 0u));
      w.meta[s] = make_uint4(m.x, m.y, depth + 1, dim);
    }
    bool valid = i < n;
    uint32_t qpos = wave_append(&w.cnt[out], valid && alive);
    if (valid && alive) w.queue[out][qpos] = s;
    uint32_t dpos = wave_append(&w.cnt[CNT_DEAD], valid && !alive);
    if (valid && !alive) w.dead[dpos] = s;
    uint32_t spos = wave_append(&w.cnt[CNT_SHADOW], emit);
    if (emit) {
      w.sh_o[spos] = so;
      w.sh_d[spos] = sdir;
      w.sh_c[spos] = sc;
    }
  }
}

struct ShadowJob {
  const float4 *sh_o, *sh_d, *sh_c;
  float4* lsum;
  __device__ __forceinline__ bool load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax, uint32_t& tag) {
    float4 a = sh_o[i], b = sh_d[i];
    o = mk(a.x, a.y, a.z);
    d = mk(b.x, b.y, b.z);
    tmin = 0.f;
    tmax = a.w;
    tag = __float_as_uint(b.w);
    return true;
  }
  __device__ __forceinline__ void finish(uint32_t i, uint32_t s, const TraceRay& r) {
    if (r.hit) return;  // occluded
    float4 c = sh_c[i];
    float4 l = lsum[s];
    lsum[s] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w);
  }
};

__global__ void __launch_bounds__(kTraceBlock) k_shadow(Wavefront w, DevScene scene) {
  ShadowJob job{w.sh_o, w.sh_d, w.sh_c, w.lsum};
  trace_stream<true>(scene, w.cnt[CNT_SHADOW], &w.cnt[CNT_FETCH_SH], job);
}

// Add the pass's samples of each pixel in sample order, dropping the ones
// the reference drops (sample_renderer.cpp:100-105: NaN anywhere; infinite
// or negative values, which the reference's HasInfinity / HasNegative look
// for in the first channel only, spectrum.h:30-31).
__global__ void k_accum(Wavefront w) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w.npix) return;
  const bool quirks = (w.flags & WF_QUIRKS) != 0;
  const size_t plane = size_t(w.rows) * w.npix;
  float4 a = w.pacc[p], a2 = w.pacc2[p];
  for (uint32_t r = 0; r < w.rows; r++) {
    size_t k = size_t(r) * w.npix + p;
    float x = w.res[k], y = w.res[k + plane], z = w.res[k + 2 * plane];
    bool bad = isnan(x) || isnan(y) || isnan(z);
    if (quirks)
      bad = bad || isinf(x) || x < 0.f;
    else
      bad = bad || isinf(x) || isinf(y) || isinf(z) || x < 0.f || y < 0.f || z < 0.f;
    if (!bad) {
      a = make_float4(a.x + x, a.y + y, a.z + z, a.w + 1.f);
      a2 = make_float4(a2.x + x * x, a2.y + y * y, a2.z + z * z, a2.w);
    }
  }
  w.pacc[p] = a;
  w.pacc2[p] = a2;
}

// fb = sum / spp (sample_renderer.cpp:109-114).
__global__ void k_final(Wavefront w, float* fb, float* sq) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w.npix) return;
  float4 a = w.pacc[p];
  float coeff = 1.0f / float(w.spp);
  fb[3 * p + 0] = a.x * coeff;
  fb[3 * p + 1] = a.y * coeff;
  fb[3 * p + 2] = a.z * coeff;
  if (sq) {
    float4 a2 = w.pacc2[p];
    sq[3 * p + 0] = a2.x;
    sq[3 * p + 1] = a2.y;
    sq[3 * p + 2] = a2.z;
  }
}

struct ClosestJob {
  const float4* rays;
  float* hits;
  __device__ __forceinline__ bool load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax, uint32_t& tag) {
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    o = mk(a.x, a.y, a.z);
    d = mk(b.x, b.y, b.z);
    tmin = a.w;
    tmax = b.w;
    tag = __float_as_uint(b.w);
    return true;
  }
  __device__ __forceinline__ void finish(uint32_t i, uint32_t tmax_bits, const TraceRay& r) {
    float* o = hits + 5 * size_t(i);
    o[0] = r.hit ? r.h.t : __uint_as_float(tmax_bits);
    o[1] = r.hit ? r.h.u : 0.f;
    o[2] = r.hit ? r.h.v : 0.f;
    o[3] = __uint_as_float(r.hit ? r.h.prim : 0xffffffffu);
    o[4] = __uint_as_float(r.hit ? r.h.geom : 0xffffffffu);
  }
};

struct AnyJob {
  const float4* rays;
  uint8_t* occ;
  __device__ __forceinline__ bool load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax, uint32_t& tag) {
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    o = mk(a.x, a.y, a.z);
    d = mk(b.x, b.y, b.z);
    tmin = a.w;
    tmax = b.w;
    tag = 0;
    return true;
  }
  __device__ __forceinline__ void finish(uint32_t i, uint32_t, const TraceRay& r) { occ[i] = r.hit ? 1 : 0; }
};

__global__ void __launch_bounds__(kTraceBlock) k_closest(DevScene scene, const float4* rays, uint32_t n, float* hits) {
  ClosestJob job{rays, hits};
  trace_stream<false>(scene, n, nullptr, job);
}

__global__ void __launch_bounds__(kTraceBlock) k_any(DevScene scene, const float4* rays, uint32_t n, uint8_t* occ) {
  AnyJob job{rays, occ};
  trace_stream<true>(scene, n, nullptr, job);
}

__global__ void k_bsdf_eval(BsdfRec b, bool quirks, const float* in, uint32_t n, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* q = in + 6 * size_t(i);
  float pdf;
  f3 f = bsdf_eval(b, quirks, mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), pdf);
  float* o = out + 4 * size_t(i);
  o[0] = f.x;
  o[1] = f.y;
  o[2] = f.z;
  o[3] = pdf;
}

__global__ void k_bsdf_sample(BsdfRec b, bool quirks, const float* in, uint32_t n, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* q = in + 6 * size_t(i);
  BsdfSample bs;
  bs.wo = splat(0.f);
  bs.eta = 1.f;
  f3 f = bsdf_sample(b, quirks, mk(q[0], q[1], q[2]), q[3], q[4], q[5], bs);
  float* o = out + 8 * size_t(i);
  o[0] = bs.wo.x;
  o[1] = bs.wo.y;
  o[2] = bs.wo.z;
  o[3] = bs.pdf;
  o[4] = f.x;
  o[5] = f.y;
  o[6] = f.z;
  o[7] = bs.eta;
}

static inline unsigned Blocks(uint64_t n, int t) { return unsigned((n + t - 1) / t); }

void LaunchInitPass(const Wavefront& w, bool first, hipStream_t s) {
  uint32_t n = first ? std::max(w.Q, w.npix) : w.Q;
  hipLaunchKernelGGL(k_init_pass, dim3(Blocks(n, 256)), dim3(256), 0, s, w, first ? 1 : 0);
}
void LaunchRegen(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_regen, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}
void LaunchReset(const Wavefront& w, int out, hipStream_t s) { hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, w.cnt, out); }
void LaunchExtend(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_extend, dim3(c.blocks), dim3(kTraceBlock), 0, s, w, scene, in);
}
void LaunchShadePath(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_shade_path, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}
void LaunchShadow(const Wavefront& w, const DevScene& scene, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_shadow, dim3(c.blocks), dim3(kTraceBlock), 0, s, w, scene);
}
void LaunchAccum(const Wavefront& w, hipStream_t s) {
  hipLaunchKernelGGL(k_accum, dim3(Blocks(w.npix, 256)), dim3(256), 0, s, w);
}
void LaunchFinal(const Wavefront& w, float* fb, float* sq, hipStream_t s) {
  hipLaunchKernelGGL(k_final, dim3(Blocks(w.npix, 256)), dim3(256), 0, s, w, fb, sq);
}
void LaunchClosest(const DevScene& scene, const void* rays, uint32_t n, void* hits, hipStream_t s) {
  if (n)
    hipLaunchKernelGGL(k_closest, dim3(Blocks(n, kTraceBlock)), dim3(kTraceBlock), 0, s, scene, (const float4*)rays, n,
                       (float*)hits);
}
void LaunchAny(const DevScene& scene, const void* rays, uint32_t n, uint8_t* occ, hipStream_t s) {
  if (n)
    hipLaunchKernelGGL(k_any, dim3(Blocks(n, kTraceBlock)), dim3(kTraceBlock), 0, s, scene, (const float4*)rays, n, occ);
}
void LaunchBsdfEval(const BsdfRec& b, bool quirks, const float* in, uint32_t n, float* out, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_bsdf_eval, dim3(Blocks(n, 128)), dim3(128), 0, s, b, quirks, in, n, out);
}
void LaunchBsdfSample(const BsdfRec& b, bool quirks, const float* in, uint32_t n, float* out, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_bsdf_sample, dim3(Blocks(n, 128)), dim3(128), 0, s, b, quirks, in, n, out);
}

}  // namespace rad
