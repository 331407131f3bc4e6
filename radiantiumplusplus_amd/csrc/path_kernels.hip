// Wavefront kernels of the `path` integrator (renderer/path.cpp:131-227) and
// the ray-batch intersection kernels (accel/embree.cpp:79-111).
#include <algorithm>

#include "guide_dev.h"
#include "guide_mlp.h"
#include "shading_dev.h"
#include "texture_dev.h"
#include "wavefront.h"

namespace rad {

// Start of a pass: Q entries on the dead list, so the first regen starts Q
// samples (that regen is told that no entry holds a finished sample, so the
// path records need no clearing here).
__global__ void k_init_pass(Wavefront w, int first) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (first && s < w.npix) {
    w.pacc[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    w.pacc2[s] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (s < w.Q) w.dead[s] = s;  // read only as a count: the first regen of a pass is told so
  if (s < 2 * w.Q) {
    if (w.glast) w.glast[s] = w.gpend[s] = 0xffffffffu;
    if (w.susp) w.susp[s] = kNoSusp;
  }
  if (s == 0) {
    w.cnt[CNT_Q0] = 0;
    w.cnt[CNT_Q1] = 0;
    w.cnt[CNT_DEAD] = w.Q;
    w.cnt[CNT_SHADOW] = 0;
    w.cnt[CNT_NEXT] = 0;
    w.cnt[CNT_FETCH_EXT] = 0;
    w.cnt[CNT_FETCH_SH] = 0;
    w.cnt[CNT_SUSP] = 0;
    w.cnt[CNT_SUSP + 1] = 0;
    if (first)
      for (int i = 0; i < ST_WORDS; i++) w.stats[i] = 0ull;
  }
}

// Training targets of a finished path: each record's weight is the
// luminance of (L_final - L_before_bounce) / throughput_before_bounce, the
// one-sample estimate of F / p for the direction sampled there (F: incident
// radiance times BSDF cosine).  A record still pending (the path ended right
// after its bounce) gets no contribution.
__device__ __forceinline__ void finish_records(const Wavefront& w, uint32_t s, f3 L) {
  uint32_t r = w.glast[s];
  const uint32_t pend = w.gpend[s];
  float wsum = 0.f;
  for (int guard = 0; r != 0xffffffffu && guard < 4096; guard++) {
    guide::TrainRecord& rc = w.recs[r];
    f3 lb = r == pend ? L : mk(rc.lbase[0], rc.lbase[1], rc.lbase[2]);
    f3 d = mk(fmaxf(L.x - lb.x, 0.f), fmaxf(L.y - lb.y, 0.f), fmaxf(L.z - lb.z, 0.f));
    f3 q = mk(rc.thr[0] > 0.f ? d.x / rc.thr[0] : 0.f, rc.thr[1] > 0.f ? d.y / rc.thr[1] : 0.f,
              rc.thr[2] > 0.f ? d.z / rc.thr[2] : 0.f);
    float wv = 0.212671f * q.x + 0.715160f * q.y + 0.072169f * q.z;
    wv = isfinite(wv) ? wv : 0.f;
    rc.w = wv;
    wsum += wv;
    r = rc.prev;
  }
  if (wsum > 0.f) atomicAdd(reinterpret_cast<float*>(&w.gcnt[G_WSUM]), wsum);
  w.glast[s] = w.gpend[s] = 0xffffffffu;
}

// Store the samples that ended last iteration into the result entries and
// start one new camera sample of the pass for each (sample_renderer.cpp:97-99),
// appended to this iteration's extend queue with its state at its position
// there.  Each thread takes kRegenPerThread dead-list entries per block
// iteration (entries base + e * 256 + thread), so one set of block atomics
// covers them all.
constexpr int kRegenPerThread = 4;

// fresh: the first regen of a pass, whose dead-list entries hold no finished sample.
__global__ void __launch_bounds__(256) k_regen(Wavefront w, DevScene scene, int in, int fresh) {
  constexpr int E = kRegenPerThread;
  __shared__ uint32_t lds[32];
  const uint32_t n = w.cnt[CNT_DEAD];
  const uint32_t stride = gridDim.x * blockDim.x * E;
  const uint32_t pass = w.rows * w.npix;
  const bool more = w.cnt[CNT_NEXT] < pass;  // stale is fine: only skips atomics once exhausted
  const uint32_t W = uint32_t(scene.cam.width);
  for (uint32_t base = blockIdx.x * blockDim.x * E; base < n; base += stride) {
    uint32_t s[E];
    int cls[E];
    bool none[E] = {};
    uint32_t finished = 0;
#pragma unroll
    for (int e = 0; e < E; e++) {
      const uint32_t i = base + e * blockDim.x + threadIdx.x;
      s[e] = 0;
      cls[e] = -1;
      if (i < n) {
        const uint32_t d = w.dead[i];
        s[e] = d & ~kDeadDone;
        cls[e] = 0;
        if (d & kDeadDone) finished++;  // result written by the shading kernel
        const uint4 m = fresh || (d & kDeadDone) ? make_uint4(0u, SAMPLE_NONE, 0u, 0u) : w.meta[s[e]];
        if (m.y != SAMPLE_NONE) {
          const float4 l = w.lsum[s[e]];
          // one 16-byte entry per sample: a single line written per finished sample
          w.res[size_t(m.y - w.row0) * w.npix + m.x] = make_float4(l.x, l.y, l.z, 0.f);
          finished++;
          if (w.flags & WF_RECORD) finish_records(w, s[e], mk(l.x, l.y, l.z));
        }
      }
    }
    uint32_t id[E];
#pragma unroll
    for (int e = 0; e < E; e++) id[e] = pass;
    if (more) {
      uint32_t* c1[1] = {&w.cnt[CNT_NEXT]};
      uint32_t pos[E], unused[E];
      block_append_n<1, E, false>(c1, cls, none, pos, unused, lds);
#pragma unroll
      for (int e = 0; e < E; e++)
        if (cls[e] >= 0) id[e] = pos[e];
    }
    int qcls[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      const bool start = cls[e] >= 0 && id[e] < pass;
      qcls[e] = start ? 0 : -1;
    }
    {
      uint32_t* c1[1] = {&w.cnt[in]};
      uint32_t pos[E], unused[E];
      block_append_n<1, E, false>(c1, qcls, none, pos, unused, lds);
#pragma unroll
      for (int e = 0; e < E; e++) {
        if (qcls[e] < 0) continue;
        // the new path's state at its queue position
        const uint32_t st = state_index(w.Q, in, pos[e]);
        const uint32_t pixel = id[e] % w.npix, sample = w.row0 + id[e] / w.npix;
        const uint32_t key = sample_key(w.seed, w.request, pixel, sample);
        w.thr[st] = make_float4(1.f, 1.f, 1.f, 1.f);
        w.lsum[st] = make_float4(0.f, 0.f, 0.f, 1.f);
        w.prev[st] = make_float4(0.f, 0.f, 0.f, __uint_as_float(PATH_SPECULAR));
        w.meta[st] = make_uint4(pixel, sample, 0u, 2u);
        w.key[st] = key;
        if (w.flags & WF_RECORD) w.glast[st] = w.gpend[st] = 0xffffffffu;
        if (w.medium) w.medium[st] = 0xffffffffu;  // vol_path: medium / channel chosen at shade time
        if (w.susp) w.susp[st] = kNoSusp;
        const uint32_t px = pixel % W, py = pixel / W;
        const float jx = rng_u(key, 0), jy = rng_u(key, 1);
        f3 o, d;
        float tmin, tmax;
        camera_ray(scene.cam, float(px) + jx, float(py) + jy, o, d, tmin, tmax);
        w.ray[in][2 * size_t(pos[e])] = make_float4(o.x, o.y, o.z, tmin);
        w.ray[in][2 * size_t(pos[e]) + 1] = make_float4(d.x, d.y, d.z, tmax);
      }
    }
    block_count(&w.stats[ST_SAMPLES], finished, lds);
  }
}

__global__ void k_reset(uint32_t* cnt, uint32_t* gcnt, int out) {
  if (threadIdx.x == 0) {
    cnt[out] = 0;
    cnt[CNT_DEAD] = 0;
    cnt[CNT_SHADOW] = 0;
    cnt[CNT_FETCH_EXT] = 0;
    cnt[CNT_FETCH_SH] = 0;
    for (int k = 0; k < kMatQueues; k++) cnt[CNT_MAT + k] = cnt[CNT_GMAT + k] = 0;
    cnt[CNT_SUSP + out] = 0;  // resume pool `out` was drained by the previous extend
    if (gcnt) gcnt[G_QUEUE] = 0;
  }
}

// The tag of an extend ray is its queue position; its state index is base + tag.
struct ExtendJob {
  static constexpr bool kSuspend = true;
  static constexpr bool kTagIsIndex = true;
  uint32_t base;
  const float4* ray;
  float4* hit;
  uint32_t* hit_geom;
  uint32_t* susp;                 // null: traversals run to the end
  const ResumeState* rin;         // pool this launch resumes from
  ResumeState* rout;              // pool this launch suspends into
  uint32_t* rcount;               // entries of rout
  uint32_t cap, budget;
  __device__ __forceinline__ void resume(uint32_t i, TraceRay& r, TraceStack& st) {
    if (!susp) return;
    const uint32_t s = base + i;
    const uint32_t j = susp[s];
    if (j == kNoSusp) return;
    restore_trace(rin[j], r, st);
    susp[s] = kNoSusp;
  }
  __device__ __forceinline__ bool suspend(uint32_t i, const TraceRay& r, const TraceStack& st) {
    const uint32_t j = atomicAdd(rcount, 1u);
    if (j >= cap) return false;  // pool full: the ray finishes in this launch
    save_trace(rout[j], r, st);
    susp[base + i] = j;
    return true;
  }
  __device__ __forceinline__ bool load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax, uint32_t& tag) {
    float4 a = ray[2 * size_t(i)], b = ray[2 * size_t(i) + 1];
    o = mk(a.x, a.y, a.z);
    d = mk(b.x, b.y, b.z);
    tmin = a.w;
    tmax = b.w;
    tag = i;
    return true;
  }
  // hits are stored by extend-queue position: consecutive rays of the
  // queue write consecutive entries, and classify reads them in order
  __device__ __forceinline__ void finish(uint32_t i, uint32_t, const TraceRay& r) {
    hit[i] = make_float4(r.h.t, r.h.u, r.h.v, __uint_as_float(r.hit ? r.h.prim : 0u));
    hit_geom[i] = r.hit ? r.h.geom : 0xffffffffu;
  }
};

// Capped at 80 VGPRs (6 waves per SIMD; four registers spill): traversal
// waits on node and primitive loads, and resident waves hide that wait, but
// the pooled primitive tests need registers; per 1080p large-mesh frame
// 78.2 ms at 6 waves, 84.8 ms at 5.
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(6))) k_extend(Wavefront w, DevScene scene, int in) {
  ExtendJob job{state_index(w.Q, in, 0), w.ray[in], w.hit, w.hit_geom, w.susp, w.resume[in], w.resume[1 - in],
                &w.cnt[CNT_SUSP + 1 - in], w.resume_cap, w.susp ? w.visit_budget : 0xffffffffu};
  trace_stream<false>(scene, w.cnt[in], &w.cnt[CNT_FETCH_EXT], job);
}

// A vertex is guided when it is a surface hit at depth <= max_depth, on a
// shape that is no emitter, whose BSDF (its record before textures) the
// guide may sample.  The one decision for the vertex: classify queues it,
// inference evaluates the network for it, and shading uses the mixture.
__device__ __forceinline__ bool guided_vertex(const Wavefront& w, const DevScene& scene, uint32_t geom, uint32_t depth,
                                              bool quirks) {
  if (w.net.max_depth != 0xffffffffu && depth > w.net.max_depth) return false;
  const ShapeRec& sr = scene.shapes[geom];
  return sr.light < 0 && sr.bsdf >= 0 && guide::guidable(scene.bsdfs[sr.bsdf], quirks, w.net.min_rough);
}

// One bounce of Path::Li (path.cpp:141-225) for every queued path.  GUIDE:
// at guided vertices the direction is drawn from the learned vMF mixture
// (probability alpha) or the BSDF, and weighted by the one-sample MIS
// density alpha q + (1 - alpha) p_bsdf; NEE uses the same density in its MIS
// weight; training records are kept for the vertices at 1 in rec_stride state indices.
// Sorted (MAT >= 0), the guided kernels shade the material's guided vertices
// and the plain ones the rest, so unguided vertices never pay for the
// mixture's registers and LDS; unsorted (MAT < 0), one kernel shades all.
template <bool GUIDE, int MAT = -1, bool TX = true>
__global__ void __launch_bounds__(256) k_shade_path(Wavefront w, DevScene scene, int in) {
  __shared__ uint32_t lds[32];
  // decoded guide mixtures, one per lane (48 KB: three blocks per CU)
  __shared__ float4 gkm[GUIDE ? guide::kLobes * 256 : 1];
  __shared__ float2 gpw[GUIDE ? guide::kLobes * 256 : 1];
  constexpr bool kSplit = GUIDE && MAT >= 0;  // a guided material queue, read from the back
  const uint32_t n = MAT >= 0 ? w.cnt[(GUIDE ? CNT_GMAT : CNT_MAT) + MAT] : w.cnt[in];
  const uint32_t* const queue = MAT >= 0 ? w.mq[MAT >= 0 ? MAT : 0] : nullptr;
  const uint32_t stride = gridDim.x * blockDim.x;
  const bool quirks = (w.flags & WF_QUIRKS) != 0;
  const bool guideOn = GUIDE && (w.flags & WF_GUIDE) != 0;
  const bool record = (w.flags & WF_RECORD) != 0;  // unguided vertices close their path's pending record too
  const float alpha = w.net.alpha;
  const int out = 1 - in;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    bool alive = false, emit = false, want_rec = false, done = false;
    uint32_t s = 0, p = 0;  // state index, and position in the extend queue (which indexes hits and rays)
    float4 so = make_float4(0, 0, 0, 0), sdir = make_float4(0, 0, 0, 0), sc = make_float4(0, 0, 0, 0);
    // the continuing path's ray and state, written at its next-queue position
    float4 nr0 = make_float4(0, 0, 0, 0), nr1 = make_float4(0, 0, 0, 0);
    float4 nthr = nr0, nlsum = nr0, nprev = nr0;
    uint4 nmeta = make_uint4(0u, 0u, 0u, 0u);
    uint32_t nkey = 0, nsusp = kNoSusp, glast = 0xffffffffu, gpend = 0xffffffffu;
    guide::TrainRecord rec;
    if (i < n) {
      p = MAT < 0 ? i : kSplit ? queue[w.Q - 1 - i] : queue[i];
      s = state_index(w.Q, in, p);
    }
    // a traversal suspended by this iteration's extend continues next iteration
    // (the material queues never hold one): its state moves to the next queue
    const bool held = MAT < 0 && i < n && w.susp && w.susp[s] != kNoSusp;
    if (held) {
      alive = true;
      nr0 = w.ray[in][2 * size_t(p)];
      nr1 = w.ray[in][2 * size_t(p) + 1];
      nthr = w.thr[s];
      nlsum = w.lsum[s];
      nprev = w.prev[s];
      nmeta = w.meta[s];
      nkey = w.key[s];
      nsusp = w.susp[s];
      if (w.glast) glast = w.glast[s], gpend = w.gpend[s];
    }
    if (i < n && !held) {
      float4 ro4 = w.ray[in][2 * size_t(p)], rd4 = w.ray[in][2 * size_t(p) + 1];
      f3 ro = mk(ro4.x, ro4.y, ro4.z), rd = mk(rd4.x, rd4.y, rd4.z);
      float4 hh = w.hit[p];
      uint32_t geom = w.hit_geom[p];
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
      if (record) {  // the previous vertex's record: radiance gathered before its bounce
        glast = w.glast[s];
        uint32_t pr = w.gpend[s];
        if (pr != 0xffffffffu) {
          w.recs[pr].lbase[0] = L.x;
          w.recs[pr].lbase[1] = L.y;
          w.recs[pr].lbase[2] = L.z;
          w.gpend[s] = 0xffffffffu;
        }
      }
      if (geom == 0xffffffffu) {
        // miss: environment emission (scene.cpp:105-113, skybox.cpp:53-63)
        if (scene.env_light >= 0) {
          float wgt = 1.f;
          if (!spec) wgt = mis_weight(prevPdf, light_pdf_dir<TX>(scene, scene.env_light, prevP, rd, rd, rd, 0.f, 0.f, 0.f, quirks));
          L += mul(T, skybox_radiance<TX>(scene, scene.lights[scene.env_light], rd)) * wgt;
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
            wgt = mis_weight(prevPdf, light_pdf_dir<TX>(scene, sr.light, prevP, si.p, si.sh.n, dir, dist, hh.y * 0.5f + 0.5f, hh.z * 0.5f + 0.5f, quirks));
          }
          f3 le = dot(si.sh.to_world(si.wi), si.n) >= 0.f
                    ? emitter_radiance<TX>(scene, scene.lights[sr.light], hh.y * 0.5f + 0.5f, hh.z * 0.5f + 0.5f)
                    : splat(0.f);
          L += mul(T, le) * wgt;
        }
        // a shape with neither light nor BSDF ends the path (the reference dereferences a null BSDF there)
        bool stop = (w.max_depth >= 0 && int32_t(depth) + 1 >= w.max_depth) || sr.light >= 0 || sr.bsdf < 0;
        if (!stop) {
          const BsdfRec* all = scene.bsdfs;
          const BsdfRec* bp0 = &scene.bsdfs[sr.bsdf];
          BsdfRec chain[3];
          if constexpr (TX && (MAT < 0 || MAT == kMatOther)) {  // textured BSDFs are classified into the generic queue
            if (bp0->tex_any) {
              // only the camera ray carries differentials (path.cpp:172; SpawnRay drops them)
              const uint32_t W = uint32_t(scene.cam.width);
              const float sx = float(m.x % W) + rng_u(key, 0), sy = float(m.x / W) + rng_u(key, 1);
              const TexCoord tc = hit_texcoord(scene, si, hh, geom, depth == 0, sx, sy);
              resolve_chain(scene, *bp0, tc, chain);
              all = chain;
              bp0 = &chain[0];
            }
          }
          const BsdfRec& b = *bp0;
          const bool guided = kSplit || (GUIDE && guided_vertex(w, scene, geom, depth, quirks));
          const guide::Mixture mx = guide::lane_mixture(gkm, gpw, threadIdx.x, 256);
          if (guided && guideOn)
            guide::decode(mx, w.graw + size_t(kSplit ? guided_row(w.cnt, MAT, i) : p) * guide::kOut, si.p, mk(w.net.center[0], w.net.center[1], w.net.center[2]),
                          mk(w.net.half[0], w.net.half[1], w.net.half[2]));
          const float a = (guided && guideOn) ? alpha : 0.f;
          if (b.lobes & LOBE_NO_DELTA) {  // NEE (path.cpp:173-191)
            float ul = rng_u(key, dim), u1 = rng_u(key, dim + 1), u2 = rng_u(key, dim + 2);
            dim += 3;
            LightSample ls = sample_light<TX>(scene, si.p, ul, u1, u2, quirks);
            if (ls.pdf > 0.f) {
              f3 wo = si.sh.to_local(ls.dir);
              float bp;
              f3 f = bsdf_eval_t<MAT>(all, b, quirks, si.wi, wo, bp);
              if (a > 0.f) bp = a * guide::mixture_pdf(mx, ls.dir) + (1.f - a) * bp;
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
          // BSDF sampling (path.cpp:192-210), or the guide with probability a
          float ul = rng_u(key, dim), u1 = rng_u(key, dim + 1), u2 = rng_u(key, dim + 2);
          dim += 3;
          BsdfSample bs;
          f3 f;
          f3 wo;
          float pdf;
          bool useGuide = false;
          if (a > 0.f) {
            useGuide = rng_u(key, dim) < a;
            dim++;
          }
          if (useGuide) {
            f3 wg = guide::mixture_sample(mx, ul, u1, u2);
            float bp;
            f = bsdf_eval_t<MAT>(all, b, quirks, si.wi, si.sh.to_local(wg), bp);
            wo = wg;
            pdf = a * guide::mixture_pdf(mx, wg) + (1.f - a) * bp;
            bs.eta = 1.f;
            bs.lobes = LOBE_GLOSSY;
            if (is_zero(f)) pdf = 0.f;  // zero-valued paths end here without bias
          } else {
            f = bsdf_sample_t<MAT>(all, b, quirks, si.wi, ul, u1, u2, bs);
            wo = si.sh.to_world(bs.wo);
            pdf = bs.pdf;
            if (a > 0.f && pdf > 0.f) pdf = a * guide::mixture_pdf(mx, wo) + (1.f - a) * bs.pdf;
          }
          if (pdf > 0.f) {
            if (record && guided && (s % w.rec_stride) == 0) {
              want_rec = true;
              f3 gn = guide::geometric_normal(scene, si.p, __float_as_uint(hh.w), geom);
              rec.p[0] = si.p.x, rec.p[1] = si.p.y, rec.p[2] = si.p.z;
              rec.n[0] = gn.x, rec.n[1] = gn.y, rec.n[2] = gn.z;
              rec.wi[0] = -rd.x, rec.wi[1] = -rd.y, rec.wi[2] = -rd.z;
              rec.dir[0] = wo.x, rec.dir[1] = wo.y, rec.dir[2] = wo.z;
              rec.thr[0] = T.x, rec.thr[1] = T.y, rec.thr[2] = T.z;
              rec.lbase[0] = rec.lbase[1] = rec.lbase[2] = 0.f;
              rec.w = 0.f;
              rec.prev = w.glast[s];
            }
            T = mul(T, f) / pdf;
            eta *= bs.eta;
            f3 no = offset_p(si.p, si.n, wo);
            nr0 = make_float4(no.x, no.y, no.z, 0.f);
            nr1 = make_float4(wo.x, wo.y, wo.z, kFloatMax);
            prevPdf = pdf;
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
      // A sample that ends here with no shadow ray still to add to it (and no
      // training records to close) is final: its result entry is written now,
      // and regen reads nothing for the path.  Of an ended sample regen reads
      // only the radiance and the pixel and sample, which stay as they are.
      if (alive) {
        nthr = make_float4(T.x, T.y, T.z, eta);
        nlsum = make_float4(L.x, L.y, L.z, prevPdf);
        nprev = make_float4(prevP.x, prevP.y, prevP.z, __uint_as_float(spec ? PATH_SPECULAR : 0u));
        nmeta = make_uint4(m.x, m.y, depth + 1, dim);
        nkey = key;
      } else if (!emit && !record) {
        w.res[size_t(m.y - w.row0) * w.npix + m.x] = make_float4(L.x, L.y, L.z, 0.f);
        done = true;
      } else {
        w.lsum[s] = make_float4(L.x, L.y, L.z, prevPdf);  // in place: the shadow ray and regen finish it
      }
    }
    bool valid = i < n;
    if (GUIDE && record) {
      uint32_t* c1[1] = {&w.gcnt[G_RECORDS]};
      bool p1[1] = {want_rec};
      uint32_t r1[1];
      block_append<1>(c1, p1, r1, lds);
      const uint32_t r = r1[0];
      if (want_rec && r < w.rec_cap) {
        w.recs[r] = rec;
        glast = gpend = r;  // baseline filled by the next bounce (or the path's end)
        if (!alive) w.glast[s] = w.gpend[s] = r;
      }
    }
    uint32_t* c3[3] = {&w.cnt[out], &w.cnt[CNT_DEAD], &w.cnt[CNT_SHADOW]};
    bool p3[3] = {valid && alive, valid && !alive, emit};
    uint32_t r3[3];
    block_append<3>(c3, p3, r3, lds);
    if (valid && alive) {
      const uint32_t d = state_index(w.Q, out, r3[0]);
      w.ray[out][2 * size_t(r3[0])] = nr0;
      w.ray[out][2 * size_t(r3[0]) + 1] = nr1;
      w.thr[d] = nthr;
      w.lsum[d] = nlsum;
      w.prev[d] = nprev;
      w.meta[d] = nmeta;
      w.key[d] = nkey;
      if (w.susp) w.susp[d] = nsusp;
      if (record || (held && w.glast)) w.glast[d] = glast, w.gpend[d] = gpend;
      sdir.w = __uint_as_float(d);  // this vertex's shadow ray adds to the moved radiance
    }
    if (valid && !alive) w.dead[r3[1]] = s | (done ? kDeadDone : 0u);
    if (emit) {
      w.sh_o[r3[2]] = so;
      w.sh_d[r3[2]] = sdir;
      w.sh_c[r3[2]] = sc;
    }
  }
}

__device__ __forceinline__ bool guided_hit(const Wavefront& w, const DevScene& scene, uint32_t geom, bool quirks,
                                           uint32_t s) {
  return geom != 0xffffffffu && guided_vertex(w, scene, geom, w.meta[s].z, quirks);
}

// Guided vertices of this iteration -> the guide queue (their extend-queue
// positions), for the unsorted shading path.
__global__ void __launch_bounds__(256) k_guide_prep(Wavefront w, DevScene scene, int in) {
  __shared__ uint32_t lds[32];
  const uint32_t n = w.cnt[in];
  const uint32_t stride = gridDim.x * blockDim.x;
  const bool quirks = (w.flags & WF_QUIRKS) != 0;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    bool g = false;
    uint32_t s = 0;
    if (i < n) {
      s = state_index(w.Q, in, i);
      g = !(w.susp && w.susp[s] != kNoSusp) && guided_hit(w, scene, w.hit_geom[i], quirks, s);
    }
    uint32_t* c1[1] = {&w.gcnt[G_QUEUE]};
    bool p1[1] = {g};
    uint32_t r1[1];
    block_append<1>(c1, p1, r1, lds);
    if (g) w.gq[r1[0]] = i;
  }
}

// Network inference for the guided vertices: the inputs (hit position,
// geometric normal, outgoing direction) are formed from the vertex's ray and
// hit, and the 40 used raw outputs go to the vertex's graw row.  SORTED: the
// vertices are the guided material queues' entries in row order (entry i of
// the concatenation is row i); otherwise the guide queue, row = position.
// Both hold extend-queue positions, which index the rays and hits.  A wave takes 64
// entries at a time, one per lane, so the chain of dependent loads behind
// the inputs (queue -> position -> ray and hit -> shape -> triangle)
// is walked for 64 vertices at once; the MLP then runs as four
// 16-vertex tiles whose inputs are shuffled in from the lanes that gathered
// them.  The weights sit in LDS once per block of kInferBlock threads.
constexpr int kInferBlock = 512;
template <bool SORTED>
__global__ void __launch_bounds__(kInferBlock) k_guide_infer_hits(Wavefront w, DevScene scene, int in) {
  __shared__ __align__(16) _Float16 sW[guide::kWeights];
  __shared__ __align__(16) _Float16 sB[guide::kBiases];
  guide::load_weights_lds_frag(sW, sB, w.net.w, w.net.b);
  __syncthreads();
  uint32_t end[kMatQueues];  // SORTED: end row of each material's guided entries
  uint32_t n = 0;
  if (SORTED) {
#pragma unroll
    for (int k = 0; k < kMatQueues; k++) end[k] = n += w.cnt[CNT_GMAT + k];
  } else {
    n = w.gcnt[G_QUEUE];
  }
  const int lane = __lane_id();
  const int v = lane & 15, g = lane >> 4;
  const uint32_t wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  for (uint32_t base = (blockIdx.x * waves + wave) * 64; base < n; base += gridDim.x * waves * 64) {
    const uint32_t i = base + lane;
    float f9[guide::kFeat];
    uint32_t pos = 0;
    if (i < n) {
      if (SORTED) {
        int k = 0;
        uint32_t j = i;
#pragma unroll
        for (int q = 0; q < kMatQueues - 1; q++)
          if (i >= end[q]) k = q + 1, j = i - end[q];
        const uint32_t* q = k == 0 ? w.mq[0] : k == 1 ? w.mq[1] : k == 2 ? w.mq[2] : k == 3 ? w.mq[3] : w.mq[4];
        pos = q[w.Q - 1 - j];
      } else {
        pos = w.gq[i];
      }
      float4 ro = w.ray[in][2 * size_t(pos)], rd = w.ray[in][2 * size_t(pos) + 1], hh = w.hit[pos];
      f3 d = mk(rd.x, rd.y, rd.z);
      f3 p = d * hh.x + mk(ro.x, ro.y, ro.z);
      f3 gn = guide::geometric_normal(scene, p, __float_as_uint(hh.w), w.hit_geom[pos]);
      f9[0] = p.x, f9[1] = p.y, f9[2] = p.z;
      f9[3] = gn.x, f9[4] = gn.y, f9[5] = gn.z;
      f9[6] = -d.x, f9[7] = -d.y, f9[8] = -d.z;
    } else {
#pragma unroll
      for (int k = 0; k < guide::kFeat; k++) f9[k] = 0.f;
    }
    const guide::FeatScalars mine = guide::feat_scalars(f9, w.net.lo, w.net.inv_ext);
    for (int t16 = 0; t16 < 4 && base + 16 * t16 < n; t16++) {  // wave-uniform bound
      const int src = 16 * t16 + v;
      guide::FeatScalars fs;
#pragma unroll
      for (int k = 0; k < 7; k++) fs.s[k] = __shfl(mine.s[k], src);
      const uint32_t row = SORTED ? base + src : __shfl(pos, src);
      guide::h4 out[3];
      guide::infer_tile(sW, sB, fs, out);
      if (base + src < n) {
        _Float16* dst = w.graw + size_t(row) * guide::kOut + 4 * g;
        *reinterpret_cast<guide::h4*>(dst) = out[0];
        *reinterpret_cast<guide::h4*>(dst + 16) = out[1];
        if (g < 2) *reinterpret_cast<guide::h4*>(dst + 32) = out[2];  // outputs 40..47 are padding
      }
    }
  }
}

// Partition the extend queue by material (misses and BSDF-less emitters go with the diffuse bucket)
// so each shade kernel runs one BSDF's code.
// GUIDE: guided vertices also go to the guide queue.  Diffuse, rough metal,
// rough glass and Disney get queues of their own (misses join the diffuse
// queue); every other type shares the last queue and its run-time-dispatch
// kernel.
// Material queue of a finished extension: misses and shapes without a BSDF
// (emitters) go to the always-launched diffuse kernel, textured BSDFs to the
// generic one.
__device__ __forceinline__ int material_key(const DevScene& scene, uint32_t geom) {
  const int32_t bi = geom == 0xffffffffu ? -1 : scene.shapes[geom].bsdf;
  if (bi < 0) return BSDF_DIFFUSE;
  const BsdfRec& br = scene.bsdfs[bi];
  return br.tex_any ? kMatOther : min(br.type, kMatOther);
}

// Each thread classifies kClassifyPerThread entries per block iteration
// (entries base + e * 256 + thread): their gathers are in flight
// together and one set of block atomics covers them all.
constexpr int kClassifyPerThread = 4;

// GUIDE (guiding or training records on): guided vertices go to their
// material's guided queue (class kMatQueues + 1 + material).
template <bool GUIDE>
__global__ void __launch_bounds__(256) k_classify(Wavefront w, DevScene scene, int in) {
  constexpr int E = kClassifyPerThread;
  // five material queues, pass-through, (five guided material queues)
  constexpr int N = GUIDE ? 2 * kMatQueues + 1 : kMatQueues + 1;
  __shared__ uint32_t lds[N * 5];
  const uint32_t n = w.cnt[in];
  const uint32_t stride = gridDim.x * blockDim.x * E;
  const bool quirks = (w.flags & WF_QUIRKS) != 0;
  uint32_t* ctr[N];
#pragma unroll
  for (int k = 0; k < kMatQueues; k++) ctr[k] = &w.cnt[CNT_MAT + k];
  ctr[kMatQueues] = &w.cnt[1 - in];
  if (GUIDE) {
#pragma unroll
    for (int k = 0; k < kMatQueues; k++) ctr[kMatQueues + 1 + k] = &w.cnt[CNT_GMAT + k];
  }
  for (uint32_t base = blockIdx.x * blockDim.x * E; base < n; base += stride) {
    uint32_t s[E];
    int cls[E];
    bool none[E] = {};
#pragma unroll
    for (int e = 0; e < E; e++) {
      const uint32_t i = base + e * blockDim.x + threadIdx.x;
      s[e] = 0;
      cls[e] = -1;
      if (i < n) {
        s[e] = state_index(w.Q, in, i);
        // traversal suspended: straight to the next extend queue
        const bool held = w.susp && w.susp[s[e]] != kNoSusp;
        if (held) {
          cls[e] = kMatQueues;
        } else {
          const uint32_t geom = w.hit_geom[i];
          cls[e] = material_key(scene, geom);
          if (GUIDE && cls[e] != BSDF_ROUGH_GLASS && guided_hit(w, scene, geom, quirks, s[e])) cls[e] += kMatQueues + 1;
        }
      }
    }
    uint32_t pos[E], unused[E];
    block_append_n<N, E, false>(ctr, cls, none, pos, unused, lds);
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int c = cls[e];
      if (c >= 0) {
        const bool g = GUIDE && c > kMatQueues;  // guided: from the back of the material's queue
        const int k = g ? c - kMatQueues - 1 : c;
        uint32_t* q = k == 0 ? w.mq[0] : k == 1 ? w.mq[1] : k == 2 ? w.mq[2] : k == 3 ? w.mq[3] : w.mq[4];
        // material queues hold extend-queue positions (which index rays,
        // hits and, with the queue, path state); a passed-through path's ray
        // and state move to its position in the next extend queue
        const uint32_t i = base + e * blockDim.x + threadIdx.x;
        if (k == kMatQueues) {
          const uint32_t d = state_index(w.Q, 1 - in, pos[e]), src = s[e];
          w.ray[1 - in][2 * size_t(pos[e])] = w.ray[in][2 * size_t(i)];
          w.ray[1 - in][2 * size_t(pos[e]) + 1] = w.ray[in][2 * size_t(i) + 1];
          w.thr[d] = w.thr[src];
          w.lsum[d] = w.lsum[src];
          w.prev[d] = w.prev[src];
          w.meta[d] = w.meta[src];
          w.key[d] = w.key[src];
          w.susp[d] = w.susp[src];
          if (w.glast) w.glast[d] = w.glast[src], w.gpend[d] = w.gpend[src];
          if (w.passsrc) w.passsrc[pos[e]] = i;
        } else {
          q[g ? w.Q - 1 - pos[e] : pos[e]] = i;
        }
      }
    }
  }
}

// Queue check (RAD_RENDER_CHECK_QUEUES).  The extension queue adds `add` to
// the tally word of each of its positions, every material queue and the
// pass-through entries of the next extension queue subtract 1, and any word
// left nonzero is a path lost, duplicated or misplaced; an entry in the wrong
// material queue, or passed through without a suspended traversal, counts as
// well.
__global__ void k_queue_tally(Wavefront w, int in, uint32_t* tally, uint32_t add) {
  const uint32_t n = w.cnt[in];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    atomicAdd(&tally[i], add);
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&w.stats[ST_QCHECK], (unsigned long long)n);
}
__global__ void k_queue_drain(Wavefront w, DevScene scene, int in, uint32_t* tally) {
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t bad = 0;
  const bool split = (w.flags & (WF_GUIDE | WF_RECORD)) != 0;
  const bool quirks = (w.flags & WF_QUIRKS) != 0;
  for (int m = 0; m <= 2 * kMatQueues; m++) {
    const bool pass = m == kMatQueues;  // the suspended paths classify passed through
    const bool g = m > kMatQueues;      // a material's guided vertices, from the back of its queue
    const int k = g ? m - kMatQueues - 1 : m;
    const uint32_t n = pass ? w.cnt[1 - in] : w.cnt[(g ? CNT_GMAT : CNT_MAT) + k];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
      // the extend-queue position each entry stands for: a material entry
      // holds it; a passed-through path sits at position i of the next queue
      // and classify recorded where it came from
      const uint32_t e = pass ? w.passsrc[i] : g ? w.mq[k][w.Q - 1 - i] : w.mq[k][i];
      if (e >= w.cnt[in]) {
        bad++;
        continue;
      }
      const uint32_t s = pass ? state_index(w.Q, 1 - in, i) : state_index(w.Q, in, e);
      const bool held = w.susp && w.susp[s] != kNoSusp;
      if (pass) {
        if (!held) bad++;
      } else if (held || material_key(scene, w.hit_geom[e]) != k ||
                 g != (split && k != BSDF_ROUGH_GLASS && guided_hit(w, scene, w.hit_geom[e], quirks, s))) {
        bad++;
      }
      atomicSub(&tally[e], 1u);
    }
  }
  if (bad) atomicAdd(&w.stats[ST_QERR], (unsigned long long)bad);
}
__global__ void k_queue_verify(Wavefront w, uint32_t* tally) {
  uint32_t bad = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < w.Q; i += gridDim.x * blockDim.x) {
    if (tally[i] != 0u) {
      bad++;
      tally[i] = 0u;
    }
  }
  if (bad) atomicAdd(&w.stats[ST_QERR], (unsigned long long)bad);
}

struct ShadowJob {
  static constexpr bool kSuspend = false;
  static constexpr bool kTagIsIndex = false;
  const float4 *sh_o, *sh_d, *sh_c;
  SlotField<float4> lsum;
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

// Capped at 72 VGPRs (7 waves per SIMD; five registers spill): per 1080p
// large-mesh frame 32.1 ms at 7 waves, 33.6 ms at 6.
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(7))) k_shadow(Wavefront w, DevScene scene) {
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
  float4 a = w.pacc[p], a2 = w.pacc2[p];
  for (uint32_t r = 0; r < w.rows; r++) {
    const float4 v = w.res[size_t(r) * w.npix + p];
    const float x = v.x, y = v.y, z = v.z;
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
  static constexpr bool kSuspend = false;
  static constexpr bool kTagIsIndex = false;
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
  static constexpr bool kSuspend = false;
  static constexpr bool kTagIsIndex = false;
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
  uint32_t n = first ? std::max(2 * w.Q, w.npix) : 2 * w.Q;
  hipLaunchKernelGGL(k_init_pass, dim3(Blocks(n, 256)), dim3(256), 0, s, w, first ? 1 : 0);
}
void LaunchRegen(const Wavefront& w, const DevScene& scene, int in, bool fresh, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_regen, dim3(c.blocks), dim3(256), 0, s, w, scene, in, fresh ? 1 : 0);
}
void LaunchReset(const Wavefront& w, int out, hipStream_t s) {
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, w.cnt, w.gcnt, out);
}
void LaunchExtend(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_extend, dim3(c.trace_blocks), dim3(kTraceBlock), 0, s, w, scene, in);
}
// TX: texture code compiled in.  The four specialised material kernels need
// it only for textured emitters (textured BSDFs are queued as "others"); the
// generic and single kernels need it for textured BSDFs as well.
template <bool G, int MAT>
static void LaunchMat(bool tx, const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  if (tx)
    hipLaunchKernelGGL((k_shade_path<G, MAT, true>), dim3(c.blocks), dim3(256), 0, s, w, scene, in);
  else
    hipLaunchKernelGGL((k_shade_path<G, MAT, false>), dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}
void LaunchShadePath(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  const bool tx = scene.tex_lights || scene.tex_bsdfs;
  if (w.flags & (WF_GUIDE | WF_RECORD))
    LaunchMat<true, -1>(tx, w, scene, in, c, s);
  else
    LaunchMat<false, -1>(tx, w, scene, in, c, s);
}
// The plain kernels shade the unguided vertices of each material; with
// guiding or training records on, the guided kernels then shade the guided
// ones (rough glass is never guided: it transmits).
static void LaunchMaterials(bool guided, const Wavefront& w, const DevScene& scene, int in, uint32_t mask,
                            const LaunchCfg& c, hipStream_t s) {
  const bool tl = scene.tex_lights != 0, tx = tl || scene.tex_bsdfs;
  // the diffuse kernel also takes the misses, so it always runs
  LaunchMat<false, BSDF_DIFFUSE>(tl, w, scene, in, c, s);
  if (mask & (1u << BSDF_ROUGH_METAL)) LaunchMat<false, BSDF_ROUGH_METAL>(tl, w, scene, in, c, s);
  if (mask & (1u << BSDF_ROUGH_GLASS)) LaunchMat<false, BSDF_ROUGH_GLASS>(tl, w, scene, in, c, s);
  if (mask & (1u << BSDF_DISNEY)) LaunchMat<false, BSDF_DISNEY>(tl, w, scene, in, c, s);
  if (mask & (1u << kMatOther)) LaunchMat<false, kMatOther>(tx, w, scene, in, c, s);
  if (!guided) return;
  if (mask & (1u << BSDF_DIFFUSE)) LaunchMat<true, BSDF_DIFFUSE>(tl, w, scene, in, c, s);
  if (mask & (1u << BSDF_ROUGH_METAL)) LaunchMat<true, BSDF_ROUGH_METAL>(tl, w, scene, in, c, s);
  if (mask & (1u << BSDF_DISNEY)) LaunchMat<true, BSDF_DISNEY>(tl, w, scene, in, c, s);
  if (mask & (1u << kMatOther)) LaunchMat<true, kMatOther>(tx, w, scene, in, c, s);
}

void LaunchQueueCheck(const Wavefront& w, const DevScene& scene, int in, uint32_t* tally, bool selftest,
                      const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_queue_tally, dim3(c.blocks), dim3(256), 0, s, w, in, tally, selftest ? 2u : 1u);
  hipLaunchKernelGGL(k_queue_drain, dim3(c.blocks), dim3(256), 0, s, w, scene, in, tally);
  hipLaunchKernelGGL(k_queue_verify, dim3(c.blocks), dim3(256), 0, s, w, tally);
}
void LaunchClassify(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  if (w.flags & (WF_GUIDE | WF_RECORD))
    hipLaunchKernelGGL(k_classify<true>, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
  else
    hipLaunchKernelGGL(k_classify<false>, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}

void LaunchShadeMaterials(const Wavefront& w, const DevScene& scene, int in, uint32_t mask, const LaunchCfg& c,
                          hipStream_t s) {
  LaunchMaterials((w.flags & (WF_GUIDE | WF_RECORD)) != 0, w, scene, in, mask, c, s);
}

void LaunchGuidePrep(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_guide_prep, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}

void LaunchGuideInfer(const Wavefront& w, const DevScene& scene, int in, int blocks, bool sorted, hipStream_t s) {
  if (sorted)
    hipLaunchKernelGGL(k_guide_infer_hits<true>, dim3(blocks), dim3(kInferBlock), 0, s, w, scene, in);
  else
    hipLaunchKernelGGL(k_guide_infer_hits<false>, dim3(blocks), dim3(kInferBlock), 0, s, w, scene, in);
}
void LaunchShadow(const Wavefront& w, const DevScene& scene, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_shadow, dim3(c.trace_blocks), dim3(kTraceBlock), 0, s, w, scene);
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
