// Kernels of the participating-media path (renderer/vol_path.cpp) and the
// heterogeneous-grid probe.
#include <algorithm>
#include <stdexcept>

#include "medium_dev.h"
#include "scene_host.h"
#include "texture_dev.h"
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

// Shadow work items of vol_path carry the current medium (+1, 8 bits), the
// spectral channel (2 bits) and the path's rng dimension at emission (22
// bits); the transmittance estimator draws from its own stream keyed by both.
__device__ __forceinline__ uint32_t pack_vol(int32_t medium, uint32_t ch, uint32_t dim) {
  return uint32_t(medium + 1) | (ch << 8) | (dim << 10);
}
__device__ __forceinline__ uint32_t shadow_key(uint32_t key, uint32_t dim) { return mix32(key + 0x9e3779b9u * (dim + 1u)); }

// One pass of VolPath::Li's loop (vol_path.cpp:27-217) for every queued path:
// Russian roulette and the depth limit at the top of the loop, delta tracking
// on one spectral channel, null collisions (counted as bounces), real
// scattering with phase-function NEE and sampling, and the surface branch
// (emission with MIS, pass-through medium boundaries that are not counted,
// BSDF NEE and sampling).  NEE contributions are finished by k_vol_shadow.
// TX: texture code compiled in (the scene has textured emitters or BSDFs).
// MAT >= 0: the kernel shades the paths k_classify put in material queue MAT
// (the BSDF of the surface the ray hit; medium events need no BSDF), with
// that BSDF's code inlined; MAT < 0 shades the whole extend queue with every
// BSDF type.  The generic kernel needs the register file of the largest BSDF
// (one wave per SIMD); the specialised ones run several.
template <bool TX, int MAT = -1>
__global__ void __launch_bounds__(256) k_shade_vol(Wavefront w, DevScene scene, int in) {
  __shared__ uint32_t lds[32];
  const uint32_t n = MAT >= 0 ? w.cnt[CNT_MAT + MAT] : w.cnt[in];
  const uint32_t stride = gridDim.x * blockDim.x;
  const bool quirks = (w.flags & WF_QUIRKS) != 0;
  const int out = 1 - in;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    bool alive = false, emit = false, msample = false, done = false;
    uint32_t s = 0;
    float4 so = make_float4(0, 0, 0, 0), sdir = make_float4(0, 0, 0, 0), sc = make_float4(0, 0, 0, 0);
    // the continuing path's ray and state, written at its next-queue position
    float4 nr0 = make_float4(0, 0, 0, 0), nr1 = make_float4(0, 0, 0, 0);
    float4 nthr = nr0, nlsum = nr0, nprev = nr0;
    uint4 nmeta = make_uint4(0u, 0u, 0u, 0u);
    uint32_t nkey = 0, nmed = 0;
    if (i < n) {
      // material queues hold extend-queue positions, which index rays, hits
      // and (with the queue) path state
      const uint32_t p = MAT >= 0 ? w.mq[MAT][i] : i;
      s = state_index(w.Q, in, p);
      float4 ro4 = w.ray[in][2 * size_t(p)], rd4 = w.ray[in][2 * size_t(p) + 1];
      f3 ro = mk(ro4.x, ro4.y, ro4.z), rd = mk(rd4.x, rd4.y, rd4.z);
      float tmin = ro4.w, tmax = rd4.w;
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
      bool spec = (__float_as_uint(p4.w) & PATH_SPECULAR) != 0;
      uint4 m = w.meta[s];
      uint32_t depth = m.z, dim = m.w;
      const uint32_t key = w.key[s];
      uint32_t medw = w.medium[s];
      int32_t medium;
      uint32_t ch;
      if (medw == MED_UNSET) {  // SampleChannel (vol_path.cpp:219-222) and the global medium
        ch = min(uint32_t(rng_u(key, dim) * 3.f), 2u);
        dim++;
        medium = scene.global_medium;
      } else {
        ch = medw >> 24;
        medium = int32_t(medw & 0xffffffu) - 1;
      }
      bool go = true;
      if (int32_t(depth) >= w.rr_depth) {
        float rr = fminf(maxc(T) * sqr(eta), 0.95f);
        if (rng_u(key, dim) > rr)
          go = false;
        else
          T = T * (1.f / rr);
        dim++;
      }
      if (go && w.max_depth >= 0 && int32_t(depth) + 1 >= w.max_depth) go = false;
      uint32_t next_depth = depth + 1;
      if (go) {
        bool isMedium = medium >= 0, isSurface = !isMedium;
        bool isNull = false, isScatter = false, escaped = false;
        const float siT = geom != 0xffffffffu ? hh.x : kInf;
        MediumSample mei;
        mei.T = kInf;
        if (isMedium) {
          const MediumRec& M = scene.media[medium];
          // the reference shortens a homogeneous ray to mei.T before tracing;
          // a hit beyond mei.T is then ignored below, so the full-length hit
          // of the extend kernel gives the same result
          mei = medium_sample(scene, M, ro, rd, tmax, rng_u(key, dim), ch);
          dim++;
          msample = true;
          if (siT < mei.T) mei.T = kInf;
          f3 tr, pdf;
          medium_tr_pdf(mei, siT, tr, pdf);
          float pc = chan(pdf, ch);
          f3 after = mul(T, tr) / pc;
          T = pc <= 0.f ? splat(0.f) : after;
          escaped = mei.T == kInf;
          isMedium = mei.T != kInf;
          bool nullScatter = rng_u(key, dim) > chan(mei.sig_t, ch) / chan(mei.comb, ch);
          dim++;
          isNull = nullScatter && isMedium;
          isScatter = !isNull && isMedium;
          if (isNull) T = mul(T, mei.sig_n) * chan(mei.comb, ch) / chan(mei.sig_n, ch);
          if (isScatter) prevP = mei.p;
        }
        if (isNull) {  // continue along the ray from the collision
          ro = mei.p;
          alive = true;
        }
        if (isScatter) {
          const MediumRec& M = scene.media[medium];
          T = mul(T, mei.sig_s) * chan(mei.comb, ch) / chan(mei.sig_t, ch);
          f3 wi = -rd;
          float ul = rng_u(key, dim), u1 = rng_u(key, dim + 1), u2 = rng_u(key, dim + 2);
          uint32_t sdim = dim;
          dim += 3;
          LightSample ls = sample_light<TX>(scene, mei.p, ul, u1, u2, quirks);
          if (ls.pdf > 0.f) {
            float ph = phase_eval(M, wi, ls.dir);
            float wgt = ls.delta ? 1.f : mis_weight(ls.pdf, ph);
            f3 c = mul(T, ls.li) * wgt * ph / ls.pdf;
            if (!is_zero(c)) {  // MediumInteraction::SpawnRay: no offset (its normal is zero)
              emit = true;
              so = make_float4(mei.p.x, mei.p.y, mei.p.z, ls.dist);
              sdir = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, __uint_as_float(s));
              sc = make_float4(c.x, c.y, c.z, __uint_as_float(pack_vol(medium, ch, sdim)));
            }
          }
          float pp;
          f3 wo = phase_sample(M, wi, rng_u(key, dim), rng_u(key, dim + 1), pp);
          dim += 2;
          ro = mei.p;
          rd = wo;
          tmin = 0.f;
          tmax = kFloatMax;
          spec = false;
          prevPdf = pp;
          alive = true;
        }
        isSurface = isSurface || escaped;
        if (isSurface) {
          if (geom == 0xffffffffu) {  // escaped: environment emission, then stop
            if (scene.env_light >= 0) {
              float wgt = 1.f;
              if (!spec) wgt = mis_weight(prevPdf, light_pdf_dir<TX>(scene, scene.env_light, prevP, rd, rd, rd, 0.f, 0.f, 0.f, quirks));
              L += mul(T, skybox_radiance<TX>(scene, scene.lights[scene.env_light], rd)) * wgt;
            }
          } else {
            SurfaceHit si = surface_hit(scene, ro, rd, hh.x, hh.y, hh.z, __float_as_uint(hh.w), geom);
            const ShapeRec& sr = scene.shapes[geom];
            if (sr.light >= 0) {
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
            } else if (sr.bsdf < 0) {  // medium boundary without a BSDF: pass through (vol_path.cpp:171-176)
              ro = offset_p(si.p, si.n, rd);
              tmin = 0.f;
              tmax = kFloatMax;
              medium = dot(rd, si.n) > 0.f ? sr.med_out : sr.med_in;
              next_depth = depth;
              alive = true;
            } else {
              const BsdfRec* all = scene.bsdfs;
              const BsdfRec* bp0 = &scene.bsdfs[sr.bsdf];
              BsdfRec chain[3];
              if constexpr (TX) {
                if (bp0->tex_any) {  // vol_path.cpp:28 drops the camera ray's differentials
                  const TexCoord tc = hit_texcoord(scene, si, hh, geom, false, 0.f, 0.f);
                  resolve_chain(scene, *bp0, tc, chain);
                  all = chain;
                  bp0 = &chain[0];
                }
              }
              const BsdfRec& b = *bp0;
              if (b.lobes & LOBE_NO_DELTA) {  // NEE through media (vol_path.cpp:178-190)
                float ul = rng_u(key, dim), u1 = rng_u(key, dim + 1), u2 = rng_u(key, dim + 2);
                uint32_t sdim = dim;
                dim += 3;
                LightSample ls = sample_light<TX>(scene, si.p, ul, u1, u2, quirks);
                if (ls.pdf > 0.f) {
                  f3 wo = si.sh.to_local(ls.dir);
                  float bp;
                  f3 f = bsdf_eval_t<MAT>(all, b, quirks, si.wi, wo, bp);
                  float wgt = ls.delta ? 1.f : mis_weight(ls.pdf, bp);
                  f3 c = mul(mul(T, f), ls.li) * wgt / ls.pdf;
                  if (!is_zero(c)) {
                    f3 o = offset_p(si.p, si.n, ls.dir);  // SpawnRay(dsr.Dir)
                    emit = true;
                    so = make_float4(o.x, o.y, o.z, ls.dist);
                    sdir = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, __uint_as_float(s));
                    sc = make_float4(c.x, c.y, c.z, __uint_as_float(pack_vol(medium, ch, sdim)));
                  }
                }
              }
              float ul = rng_u(key, dim), u1 = rng_u(key, dim + 1), u2 = rng_u(key, dim + 2);
              dim += 3;
              BsdfSample bs;
              f3 f = bsdf_sample_t<MAT>(all, b, quirks, si.wi, ul, u1, u2, bs);
              if (bs.pdf > 0.f) {
                T = mul(T, f) / bs.pdf;
                eta *= bs.eta;
                f3 wo = si.sh.to_world(bs.wo);
                ro = offset_p(si.p, si.n, wo);
                rd = wo;
                tmin = 0.f;
                tmax = kFloatMax;
                prevPdf = bs.pdf;
                prevP = si.p;
                spec = (bs.lobes & LOBE_DELTA) != 0;
                medium = dot(wo, si.n) > 0.f ? sr.med_out : sr.med_in;
                alive = true;
              }
            }
          }
        }
      }
      if (alive) {
        nr0 = make_float4(ro.x, ro.y, ro.z, tmin);
        nr1 = make_float4(rd.x, rd.y, rd.z, tmax);
        nthr = make_float4(T.x, T.y, T.z, eta);
        nlsum = make_float4(L.x, L.y, L.z, prevPdf);
        nprev = make_float4(prevP.x, prevP.y, prevP.z, __uint_as_float(spec ? PATH_SPECULAR : 0u));
        nmeta = make_uint4(m.x, m.y, next_depth, dim);
        nmed = uint32_t(medium + 1) | (ch << 24);
        nkey = key;
      }
      // a sample that ends here with no shadow ray still to add to it is
      // final: its result entry is written now, and regen reads nothing for it
      if (!alive && !emit) {
        w.res[size_t(m.y - w.row0) * w.npix + m.x] = make_float4(L.x, L.y, L.z, 0.f);
        done = true;
      } else if (!alive) {
        w.lsum[s] = make_float4(L.x, L.y, L.z, prevPdf);  // in place: the shadow ray and regen finish it
      }
    }
    bool valid = i < n;
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
      w.medium[d] = nmed;
      w.key[d] = nkey;
      sdir.w = __uint_as_float(d);  // this vertex's shadow ray adds to the moved radiance
    }
    if (valid && !alive) w.dead[r3[1]] = s | (done ? kDeadDone : 0u);
    if (emit) {
      w.sh_o[r3[2]] = so;
      w.sh_d[r3[2]] = sdir;
      w.sh_c[r3[2]] = sc;
    }
    block_count(&w.stats[ST_MEDIUM], msample ? 1u : 0u, lds);
  }
}

// The transmittance loop of VolPath::SampleLight (vol_path.cpp:224-342):
// ratio tracking on the path's channel through media, passing through
// surfaces without a BSDF, blocked by any surface with one.  The reference
// intersects the scene again at every medium collision; here each straight
// segment is traced once and the surface distance is carried across null
// collisions (si.T - mei.T, as vol_path.cpp:300 keeps it), which changes
// nothing but the cost.  Lanes are refilled from the shadow queue as their
// rays finish, so one long transmittance walk does not hold its wave.
struct VolShadowRay {
  f3 o, d, oTrace;  // origin, direction, origin the cached hit distance h.t refers to
  f3 tr;
  float dist, totalDist, siT;
  uint32_t s, skey, sdim, ch;
  int32_t medium;
  bool needsIntersection, hitValid;
  Hit h;
  float4 c;  // contribution
};

__device__ __forceinline__ void vol_shadow_begin(const Wavefront& w, uint32_t i, VolShadowRay& r) {
  const float4 o4 = w.sh_o[i], d4 = w.sh_d[i];
  r.c = w.sh_c[i];
  r.s = __float_as_uint(d4.w);
  const uint32_t pk = __float_as_uint(r.c.w);
  r.medium = int32_t(pk & 0xffu) - 1;
  r.ch = (pk >> 8) & 3u;
  r.skey = shadow_key(w.key[r.s], pk >> 10);
  r.sdim = 0;
  r.o = mk(o4.x, o4.y, o4.z);
  r.d = mk(d4.x, d4.y, d4.z);
  r.oTrace = r.o;
  r.dist = o4.w;
  r.totalDist = 0.f;
  r.tr = splat(1.f);
  r.needsIntersection = true;
  r.hitValid = false;
  r.siT = kInf;
}

// One pass of the reference's while loop; returns false once the ray is done.
__device__ __forceinline__ bool vol_shadow_step(const DevScene& scene, VolShadowRay& r, uint32_t& steps) {
  const float remaining = r.dist * (1.f - kShadowEps) - r.totalDist;
  if (!(remaining > 0.f)) return false;
  bool escaped = false;
  bool activeMedium = r.medium >= 0;
  bool activeSurface = !activeMedium;
  // The reference intersects in the medium branch (vol_path.cpp:258-263) or
  // the surface branch (:316-320), whichever is active, from the same origin
  // and only when a new straight segment starts; one call site serves both.
  if (r.needsIntersection) {
    r.hitValid = trace<false>(scene, r.o, r.d, 0.f, remaining, r.h);
    r.siT = r.hitValid ? r.h.t : kInf;
    r.oTrace = r.o;
    r.needsIntersection = false;
  }
  if (activeMedium) {
    const MediumRec& M = scene.media[r.medium];
    MediumSample mei = medium_sample(scene, M, r.o, r.d, remaining, rng_u(r.skey, r.sdim++), r.ch);
    steps++;
    if (r.siT < mei.T) mei.T = kInf;
    float t = fminf(remaining, fminf(mei.T, r.siT)) - mei.minT;
    if (t >= 0.f) {
      f3 e = mk(expf(mei.comb.x * -t), expf(mei.comb.y * -t), expf(mei.comb.z * -t));
      f3 ffp = (r.siT < mei.T || mei.T > remaining) ? e : mul(e, mei.comb);
      float trPdf = chan(ffp, r.ch);
      f3 mTr = trPdf > 0.f ? e / trPdf : splat(0.f);
      r.tr = mul(r.tr, mTr);
    }
    if (mei.T > remaining && mei.T != kInf) r.totalDist = r.dist;
    if (mei.T > remaining) mei.T = kInf;
    escaped = mei.T == kInf;
    activeMedium = mei.T != kInf;
    if (activeMedium) {
      r.totalDist += mei.T;
      r.o = mei.p;
      r.siT = r.siT - mei.T;
      r.tr = mul(r.tr, mei.sig_n);
    }
  }
  activeSurface = activeSurface || escaped;
  if (activeSurface) r.totalDist += r.siT;
  activeSurface = activeSurface && r.hitValid && r.siT != kInf && !activeMedium;
  int32_t nextMedium = r.medium;
  if (activeSurface) {
    const ShapeRec& sr = scene.shapes[r.h.geom];
    if (sr.bsdf >= 0) r.tr = splat(0.f);
    // h.t was measured from the origin of the trace that found the hit
    SurfaceHit si = surface_hit(scene, r.oTrace, r.d, r.h.t, r.h.u, r.h.v, r.h.prim, r.h.geom);
    r.o = offset_p(si.p, si.n, r.d);
    if (sr.med_in >= 0 || sr.med_out >= 0) nextMedium = dot(r.d, si.n) > 0.f ? sr.med_out : sr.med_in;
  }
  r.needsIntersection = r.needsIntersection || activeSurface;
  r.medium = nextMedium;
  return (activeMedium || activeSurface) && !is_zero(r.tr);
}

__global__ void __launch_bounds__(kTraceBlock) k_vol_shadow(Wavefront w, DevScene scene) {
  __shared__ uint32_t lds[32];
  const uint32_t n = w.cnt[CNT_SHADOW];
  uint32_t* fetch = &w.cnt[CNT_FETCH_SH];
  constexpr uint32_t kChunk = 256;  // queue entries a wave takes with one atomic
  const uint32_t lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  VolShadowRay r;
  bool active = false, exhausted = false;
  uint32_t cbase = 0, cleft = 0, steps = 0, guard = 0;
  for (;;) {
    const unsigned long long need = __ballot(!active);
    if (need && (!exhausted || cleft > 0)) {
      const uint32_t want = (uint32_t)__popcll(need);
      const uint32_t rank = (uint32_t)__popcll(need & below);
      uint32_t nbase = 0, nleft = 0;
      if (cleft < want && !exhausted) {
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(fetch, kChunk);
        b = __shfl(b, 0);
        if (b >= n) {
          exhausted = true;
        } else {
          nbase = b;
          nleft = min(kChunk, n - b);
          if (b + kChunk >= n) exhausted = true;
        }
      }
      uint32_t i = 0xffffffffu;
      if (rank < cleft)
        i = cbase + rank;
      else if (rank - cleft < nleft)
        i = nbase + (rank - cleft);
      const uint32_t fromOld = min(want, cleft);
      cbase += fromOld;
      cleft -= fromOld;
      if (nleft) {
        const uint32_t fromNew = min(want - fromOld, nleft);
        cbase = nbase + fromNew;
        cleft = nleft - fromNew;
      }
      if (!active && i < n) {
        vol_shadow_begin(w, i, r);
        active = true;
        guard = 0;
      }
    }
    if (__ballot(active) == 0ull) {
      if (exhausted && cleft == 0) break;
      continue;
    }
    if (active) {
      // the guard bounds a walk that no longer moves (finite distances end)
      if (!vol_shadow_step(scene, r, steps) || ++guard >= 100000) {
        if (!is_zero(r.tr)) {
          const float4 l = w.lsum[r.s];
          w.lsum[r.s] = make_float4(l.x + r.c.x * r.tr.x, l.y + r.c.y * r.tr.y, l.z + r.c.z * r.tr.z, l.w);
        }
        active = false;
      }
    }
  }
  block_count(&w.stats[ST_MEDIUM], steps, lds);
}

void LaunchShadeVol(const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  if (scene.tex_lights || scene.tex_bsdfs)
    hipLaunchKernelGGL(k_shade_vol<true>, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
  else
    hipLaunchKernelGGL(k_shade_vol<false>, dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}

template <int MAT>
static void LaunchVolMat(bool tx, const Wavefront& w, const DevScene& scene, int in, const LaunchCfg& c, hipStream_t s) {
  if (tx)
    hipLaunchKernelGGL((k_shade_vol<true, MAT>), dim3(c.blocks), dim3(256), 0, s, w, scene, in);
  else
    hipLaunchKernelGGL((k_shade_vol<false, MAT>), dim3(c.blocks), dim3(256), 0, s, w, scene, in);
}

// After k_classify: one kernel per material queue present (the diffuse queue
// also holds misses and medium boundaries, so it always runs).
void LaunchShadeVolMaterials(const Wavefront& w, const DevScene& scene, int in, uint32_t mask, const LaunchCfg& c,
                             hipStream_t s) {
  const bool tl = scene.tex_lights != 0, tx = tl || scene.tex_bsdfs;
  LaunchVolMat<BSDF_DIFFUSE>(tl, w, scene, in, c, s);
  if (mask & (1u << BSDF_ROUGH_METAL)) LaunchVolMat<BSDF_ROUGH_METAL>(tl, w, scene, in, c, s);
  if (mask & (1u << BSDF_ROUGH_GLASS)) LaunchVolMat<BSDF_ROUGH_GLASS>(tl, w, scene, in, c, s);
  if (mask & (1u << BSDF_DISNEY)) LaunchVolMat<BSDF_DISNEY>(tl, w, scene, in, c, s);
  if (mask & (1u << kMatOther)) LaunchVolMat<kMatOther>(tx, w, scene, in, c, s);
}
void LaunchVolShadow(const Wavefront& w, const DevScene& scene, const LaunchCfg& c, hipStream_t s) {
  hipLaunchKernelGGL(k_vol_shadow, dim3(c.trace_blocks), dim3(kTraceBlock), 0, s, w, scene);
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
