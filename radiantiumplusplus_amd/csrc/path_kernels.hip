// Wavefront kernels of the `path` integrator (renderer/path.cpp:131-227) and
// the ray-batch intersection kernels (accel/embree.cpp:79-111).
#include "shading_dev.h"
#include "wavefront.h"

namespace rad {

// Append `pred` lanes of this wavefront to a queue with one atomic per wave.
// Must be reached by every lane of the wave (uniform control flow).
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return 0;
  uint32_t lane = __lane_id();
  uint32_t leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ void wave_count(unsigned long long* counter, bool pred) {
  unsigned long long mask = __ballot(pred);
  if (mask != 0ull && __lane_id() == (uint32_t)(__ffsll((long long)mask) - 1))
    atomicAdd(counter, (unsigned long long)__popcll(mask));
}

__global__ void k_init(Wavefront w) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= w.Q) return;
  w.meta[s] = make_uint4(s % w.npix, SAMPLE_NONE, 0u, 0u);
  w.acc[s] = make_float4(0.f, 0.f, 0.f, 0.f);
  w.acc2[s] = make_float4(0.f, 0.f, 0.f, 0.f);
  w.dead[s] = s;
  if (s == 0) {
    w.cnt[CNT_Q0] = 0;
    w.cnt[CNT_Q1] = 0;
    w.cnt[CNT_DEAD] = w.Q;
    w.cnt[CNT_SHADOW] = 0;
    for (int i = 0; i < ST_WORDS; i++) w.stats[i] = 0ull;
  }
}

// Finalise the sample a dead slot carried (sample_renderer.cpp:100-105) and
// start its next camera sample (sample_renderer.cpp:97-99).
__global__ void __launch_bounds__(256) k_regen(Wavefront w, DevScene scene, int in) {
  const uint32_t n = w.cnt[CNT_DEAD];
  const uint32_t stride = gridDim.x * blockDim.x;
  const bool quirks = (w.flags & WF_QUIRKS) != 0;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < n;
    bool start = false, finished = false;
    uint32_t s = 0;
    if (valid) {
      s = w.dead[i];
      uint4 m = w.meta[s];
      if (m.y < w.spp) {
        float4 l = w.lsum[s];
        bool bad = isnan(l.x) || isnan(l.y) || isnan(l.z);
        // the reference's HasInfinity / HasNegative look at the first channel only (spectrum.h:30-31)
        if (quirks)
          bad = bad || isinf(l.x) || l.x < 0.f;
        else
          bad = bad || isinf(l.x) || isinf(l.y) || isinf(l.z) || l.x < 0.f || l.y < 0.f || l.z < 0.f;
        if (!bad) {
          float4 a = w.acc[s], a2 = w.acc2[s];
          w.acc[s] = make_float4(a.x + l.x, a.y + l.y, a.z + l.z, a.w + 1.f);
          w.acc2[s] = make_float4(a2.x + l.x * l.x, a2.y + l.y * l.y, a2.z + l.z * l.z, a2.w);
        }
        finished = true;
      }
      uint32_t lane = s / w.npix;
      uint32_t next = m.y == SAMPLE_NONE ? lane : (m.y >= w.spp ? SAMPLE_DONE : m.y + w.K);
      if (next < w.spp) {
        uint32_t pixel = m.x;
        uint32_t key = sample_key(w.seed, w.request, pixel, next);
        uint32_t px = pixel % uint32_t(scene.cam.width), py = pixel / uint32_t(scene.cam.width);
        float jx = rng_u(key, 0), jy = rng_u(key, 1);
        f3 o, d;
        float tmin, tmax;
        camera_ray(scene.cam, float(px) + jx, float(py) + jy, o, d, tmin, tmax);
        w.ray_o[s] = make_float4(o.x, o.y, o.z, tmin);
        w.ray_d[s] = make_float4(d.x, d.y, d.z, tmax);
        w.thr[s] = make_float4(1.f, 1.f, 1.f, 1.f);
        w.lsum[s] = make_float4(0.f, 0.f, 0.f, 1.f);
        w.prev[s] = make_float4(0.f, 0.f, 0.f, __uint_as_float(PATH_SPECULAR));
        w.meta[s] = make_uint4(pixel, next, 0u, 2u);
        w.key[s] = key;
        if (w.medium) w.medium[s] = 0xffffffffu;  // vol_path: medium / channel chosen at shade time
        start = true;
      } else {
        w.meta[s] = make_uint4(m.x, SAMPLE_DONE, 0u, 0u);
      }
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
  }
}

__global__ void __launch_bounds__(256) k_extend(Wavefront w, DevScene scene, int in) {
  const uint32_t n = w.cnt[in];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    if (i < n) {
      uint32_t s = w.queue[in][i];
      float4 ro = w.ray_o[s], rd = w.ray_d[s];
      Hit h;
      h.geom = 0xffffffffu;
      bool hit = trace<false>(scene, mk(ro.x, ro.y, ro.z), mk(rd.x, rd.y, rd.z), ro.w, rd.w, h);
      w.hit[s] = make_float4(h.t, h.u, h.v, __uint_as_float(hit ? h.prim : 0u));
      w.hit_geom[s] = hit ? h.geom : 0xffffffffu;
    }
    wave_count(&w.stats[ST_SEGMENTS], i < n);
  }
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

__global__ void __launch_bounds__(256) k_shadow(Wavefront w, DevScene scene) {
  const uint32_t n = w.cnt[CNT_SHADOW];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    if (i < n) {
      float4 o = w.sh_o[i], d = w.sh_d[i];
      Hit h;
      bool occ = trace<true>(scene, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), 0.f, o.w, h);
      if (!occ) {
        uint32_t s = __float_as_uint(d.w);
        float4 c = w.sh_c[i];
        float4 l = w.lsum[s];
        w.lsum[s] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w);
      }
    }
    wave_count(&w.stats[ST_SHADOW], i < n);
  }
}

// fb = (sum over the K slots of a pixel) / spp (sample_renderer.cpp:109-114).
__global__ void k_resolve(Wavefront w, float* fb, float* sq) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w.npix) return;
  float r = 0.f, g = 0.f, b = 0.f, r2 = 0.f, g2 = 0.f, b2 = 0.f;
  for (uint32_t k = 0; k < w.K; k++) {
    float4 a = w.acc[p + k * w.npix], a2 = w.acc2[p + k * w.npix];
    r += a.x;
    g += a.y;
    b += a.z;
    r2 += a2.x;
    g2 += a2.y;
    b2 += a2.z;
  }
  float coeff = 1.0f / float(w.spp);
  fb[3 * p + 0] = r * coeff;
  fb[3 * p + 1] = g * coeff;
  fb[3 * p + 2] = b * coeff;
  if (sq) {
    sq[3 * p + 0] = r2;
    sq[3 * p + 1] = g2;
    sq[3 * p + 2] = b2;
  }
}

__global__ void __launch_bounds__(256) k_closest(DevScene scene, const float4* rays, uint32_t n, float* hits) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 a = rays[2 * i], b = rays[2 * i + 1];
  Hit h;
  bool hit = trace<false>(scene, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w, h);
  float* o = hits + 5 * size_t(i);
  o[0] = hit ? h.t : b.w;
  o[1] = hit ? h.u : 0.f;
  o[2] = hit ? h.v : 0.f;
  o[3] = __uint_as_float(hit ? h.prim : 0xffffffffu);
  o[4] = __uint_as_float(hit ? h.geom : 0xffffffffu);
}

__global__ void __launch_bounds__(256) k_any(DevScene scene, const float4* rays, uint32_t n, uint8_t* occ) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 a = rays[2 * i], b = rays[2 * i + 1];
  Hit h;
  occ[i] = trace<true>(scene, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w, h) ? 1 : 0;
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

void LaunchInit(const Wavefront& w, hipStream_t s) {
  hipLaunchKernelGGL(k_init, dim3(Blocks(w.Q, 256)), dim3(256), 0, s, w);
}
void LaunchRegen(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_regen, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}
void LaunchReset(const Wavefront& w, int out, hipStream_t s) { hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, w.cnt, out); }
void LaunchExtend(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_extend, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}
void LaunchShadePath(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_shade_path, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}
void LaunchShadow(const Wavefront& w, const DevScene& scene, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_shadow, dim3(c.blocks), dim3(256), 0, s, w, scene);
}
void LaunchResolve(const Wavefront& w, float* fb, float* sq, hipStream_t s) {
  hipLaunchKernelGGL(k_resolve, dim3(Blocks(w.npix, 256)), dim3(256), 0, s, w, fb, sq);
}
void LaunchClosest(const DevScene& scene, const void* rays, uint32_t n, void* hits, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_closest, dim3(Blocks(n, 256)), dim3(256), 0, s, scene, (const float4*)rays, n, (float*)hits);
}
void LaunchAny(const DevScene& scene, const void* rays, uint32_t n, uint8_t* occ, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_any, dim3(Blocks(n, 256)), dim3(256), 0, s, scene, (const float4*)rays, n, occ);
}
void LaunchBsdfEval(const BsdfRec& b, bool quirks, const float* in, uint32_t n, float* out, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_bsdf_eval, dim3(Blocks(n, 128)), dim3(128), 0, s, b, quirks, in, n, out);
}
void LaunchBsdfSample(const BsdfRec& b, bool quirks, const float* in, uint32_t n, float* out, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_bsdf_sample, dim3(Blocks(n, 128)), dim3(128), 0, s, b, quirks, in, n, out);
}

}  // namespace rad
