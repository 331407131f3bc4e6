#include "render_job.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>

#include "wavefront.h"

struct rad_job {
  rad_scene* scene = nullptr;
  rad_render_desc desc{};
  rad::Wavefront w{};
  rad::DeviceBuffer arena;
  rad::DeviceBuffer d_fb, d_sq;
  hipStream_t stream = nullptr;
  uint32_t* h_cnt = nullptr;  // pinned readback: counters + stats
  std::thread th;
  std::atomic<bool> cancel{false};
  std::atomic<bool> complete{false};
  std::atomic<uint64_t> done{0};
  uint64_t total = 0;
  std::string error;
  rad_status status = RAD_OK;
  double render_ms = 0;
  uint32_t iterations = 0;
  uint32_t passes = 0;
  uint64_t segments = 0, shadow_rays = 0;
  uint64_t stats[rad::ST_WORDS] = {};
  double kernel_ms[RAD_KERNEL_COUNT] = {};
  uint32_t kernel_launches[RAD_KERNEL_COUNT] = {};
  bool joined = false;
};

namespace rad {

namespace {

constexpr uint64_t kDefaultSlots = 4u << 20;     // paths in flight
constexpr uint64_t kResultBudget = 2ull << 30;   // bytes of result planes per pass
size_t Align(size_t v) { return (v + 255) & ~size_t(255); }

struct JobShape {
  uint32_t npix, rows, Q;
  uint64_t bytes;
};

JobShape Shape(const rad_scene* scene, const rad_render_desc& desc) {
  JobShape j{};
  j.npix = uint32_t(scene->host.cam.width) * uint32_t(scene->host.cam.height);
  const uint64_t spp = std::max(1u, desc.spp);
  uint64_t rows = kResultBudget / (uint64_t(j.npix) * 12);
  rows = std::min<uint64_t>(std::max<uint64_t>(rows, 1), spp);
  rows = std::min<uint64_t>(rows, 0x7fffffffull / j.npix);
  if (rows == 0) throw StatusError(RAD_ERR_INVALID_ARG, "image too large");
  j.rows = uint32_t(rows);
  const uint64_t pass = rows * j.npix;
  uint64_t Q = desc.path_slots > 0 ? desc.path_slots : kDefaultSlots;
  Q = std::max<uint64_t>(1, std::min(Q, pass));
  j.Q = uint32_t(Q);
  j.bytes = Align(Q * 16) * 10 + Align(Q * 4) * 6 + Align(pass * 4) * 3 + Align(uint64_t(j.npix) * 16) * 2 +
            2 * Align(uint64_t(j.npix) * 12) + 4096;
  return j;
}

void RunJob(rad_job* job) {
  try {
    HipCheck(hipSetDevice(job->scene->ctx->device), "hipSetDevice");
    const DevScene scene = job->scene->dev;
    Wavefront& w = job->w;
    hipStream_t st = job->stream;
    LaunchCfg cfg{std::max(64, job->scene->ctx->cu_count * 8), 256};
    const bool prof = (job->desc.flags & RAD_RENDER_PROFILE) != 0;
    hipEvent_t ev[RAD_KERNEL_COUNT + 1] = {};
    if (prof)
      for (auto& e : ev) HipCheck(hipEventCreate(&e), "hipEventCreate");
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t rows = w.rows;
    for (uint32_t row0 = 0; row0 < w.spp && job->status == RAD_OK; row0 += rows) {
      w.row0 = row0;
      w.rows = std::min(rows, w.spp - row0);
      LaunchInitPass(w, row0 == 0, st);
      int in = 0;
      for (;;) {
        if (prof) HipCheck(hipEventRecord(ev[0], st), "event");
        LaunchRegen(w, scene, in, cfg, st);
        LaunchReset(w, 1 - in, st);
        if (prof) HipCheck(hipEventRecord(ev[1], st), "event");
        LaunchExtend(w, scene, in, cfg, st);
        if (prof) HipCheck(hipEventRecord(ev[2], st), "event");
        const bool vol = w.integrator == RAD_INTEGRATOR_VOL_PATH;
        if (vol)
          LaunchShadeVol(w, scene, in, cfg, st);
        else
          LaunchShadePath(w, scene, in, cfg, st);
        if (prof) HipCheck(hipEventRecord(ev[3], st), "event");
        if (vol)
          LaunchVolShadow(w, scene, cfg, st);
        else
          LaunchShadow(w, scene, cfg, st);
        if (prof) HipCheck(hipEventRecord(ev[4], st), "event");
        HipCheck(hipGetLastError(), "wavefront launch");
        HipCheck(hipMemcpyAsync(job->h_cnt, w.cnt, CNT_WORDS * sizeof(uint32_t) + ST_WORDS * sizeof(uint64_t),
                                hipMemcpyDeviceToHost, st),
                 "counter readback");
        HipCheck(hipStreamSynchronize(st), "wavefront iteration");
        job->iterations++;
        if (prof) {
          for (int k = 0; k < 4; k++) {
            float ms = 0.f;
            HipCheck(hipEventElapsedTime(&ms, ev[k], ev[k + 1]), "event time");
            job->kernel_ms[k] += ms;
            job->kernel_launches[k]++;
          }
        }
        const uint64_t* hs = reinterpret_cast<const uint64_t*>(job->h_cnt + CNT_WORDS);
        job->done.store(hs[ST_SAMPLES]);
        job->segments += job->h_cnt[in];
        job->shadow_rays += job->h_cnt[CNT_SHADOW];
        uint32_t cont = job->h_cnt[1 - in], dead = job->h_cnt[CNT_DEAD];
        if (cont == 0 && dead == 0) break;
        if (job->iterations > 1000000u)
          throw StatusError(RAD_ERR_STATE, "path lengths exceed one million bounces (no termination criterion)");
        if (job->cancel.load()) {
          job->status = RAD_ERR_CANCELLED;
          break;
        }
        in = 1 - in;
      }
      if (job->status != RAD_OK) break;
      LaunchAccum(w, st);
      job->passes++;
    }
    w.rows = rows;
    LaunchFinal(w, job->d_fb.as<float>(), job->d_sq.as<float>(), st);
    HipCheck(hipGetLastError(), "resolve launch");
    HipCheck(hipStreamSynchronize(st), "resolve");
    job->render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    HipCheck(hipMemcpy(job->h_cnt, w.cnt, CNT_WORDS * sizeof(uint32_t) + ST_WORDS * sizeof(uint64_t),
                       hipMemcpyDeviceToHost),
             "stats readback");
    const uint64_t* hs = reinterpret_cast<const uint64_t*>(job->h_cnt + CNT_WORDS);
    for (int i = 0; i < ST_WORDS; i++) job->stats[i] = hs[i];
    if (prof)
      for (auto& e : ev) (void)hipEventDestroy(e);
  } catch (const StatusError& e) {
    job->status = e.code;
    job->error = e.what();
  } catch (const std::exception& e) {
    job->status = RAD_ERR_HIP;
    job->error = e.what();
  }
  job->complete.store(true);
}

}  // namespace

uint64_t JobFootprint(const rad_scene* scene, const rad_render_desc& desc, uint32_t* slots) {
  JobShape j = Shape(scene, desc);
  if (slots) *slots = j.Q;
  return j.bytes;
}

rad_job* StartJob(rad_scene* scene, const rad_render_desc& desc) {
  if (desc.spp == 0) throw StatusError(RAD_ERR_INVALID_ARG, "spp must be positive");
  if (desc.integrator != RAD_INTEGRATOR_PATH && desc.integrator != RAD_INTEGRATOR_VOL_PATH)
    throw StatusError(RAD_ERR_INVALID_ARG, "unknown integrator");
  HipCheck(hipSetDevice(scene->ctx->device), "hipSetDevice");
  // the reference dereferences the BSDF of every hit surface in `path` (path.cpp:172)
  if (desc.integrator == RAD_INTEGRATOR_PATH) {
    for (size_t i = 0; i < scene->host.shapes.size(); i++)
      if (scene->host.shapes[i].bsdf < 0)
        throw StatusError(RAD_ERR_INVALID_ARG, "shape " + std::to_string(i) + " has no BSDF (path needs one)");
  }
  // vol_path shadow work items carry the medium index in 8 bits
  if (desc.integrator == RAD_INTEGRATOR_VOL_PATH && scene->host.media.size() > 254)
    throw StatusError(RAD_ERR_UNSUPPORTED, "vol_path supports at most 254 media per scene");
  auto job = std::make_unique<rad_job>();
  job->scene = scene;
  job->desc = desc;
  const JobShape js = Shape(scene, desc);
  const uint32_t npix = js.npix, Q = js.Q;
  job->total = uint64_t(npix) * desc.spp;
  // one arena for every per-slot, per-sample and per-pixel array
  job->arena.alloc(js.bytes);
  char* p = static_cast<char*>(job->arena.ptr);
  auto take = [&](size_t n) {
    char* r = p;
    p += Align(n);
    return r;
  };
  Wavefront& w = job->w;
  const size_t f4 = size_t(Q) * 16, u4 = size_t(Q) * 4;
  w.ray_o = (float4*)take(f4);
  w.ray_d = (float4*)take(f4);
  w.thr = (float4*)take(f4);
  w.lsum = (float4*)take(f4);
  w.prev = (float4*)take(f4);
  w.meta = (uint4*)take(f4);
  w.hit = (float4*)take(f4);
  w.sh_o = (float4*)take(f4);
  w.sh_d = (float4*)take(f4);
  w.sh_c = (float4*)take(f4);
  w.key = (uint32_t*)take(u4);
  w.medium = (uint32_t*)take(u4);
  w.hit_geom = (uint32_t*)take(u4);
  w.queue[0] = (uint32_t*)take(u4);
  w.queue[1] = (uint32_t*)take(u4);
  w.dead = (uint32_t*)take(u4);
  w.res = (float*)take(size_t(js.rows) * npix * 4 * 3);
  w.pacc = (float4*)take(size_t(npix) * 16);
  w.pacc2 = (float4*)take(size_t(npix) * 16);
  w.cnt = (uint32_t*)take(CNT_WORDS * 4 + ST_WORDS * 8);
  w.stats = reinterpret_cast<unsigned long long*>(w.cnt + CNT_WORDS);
  job->d_sq.alloc(size_t(npix) * 3 * sizeof(float));
  job->d_fb.alloc(size_t(npix) * 3 * sizeof(float));
  w.npix = npix;
  w.Q = Q;
  w.spp = desc.spp;
  w.row0 = 0;
  w.rows = js.rows;
  w.seed = desc.seed;
  w.request = desc.request_id;
  w.max_depth = desc.max_depth;
  w.rr_depth = desc.rr_depth;
  w.flags = (desc.flags & RAD_RENDER_FIX_QUIRKS) ? 0u : WF_QUIRKS;
  w.integrator = desc.integrator;
  HipCheck(hipHostMalloc((void**)&job->h_cnt, CNT_WORDS * 4 + ST_WORDS * 8, hipHostMallocDefault), "hipHostMalloc");
  HipCheck(hipStreamCreateWithFlags(&job->stream, hipStreamNonBlocking), "hipStreamCreate");
  rad_job* raw = job.release();
  raw->th = std::thread(RunJob, raw);
  return raw;
}

void PollJob(rad_job* job, uint64_t* done, uint64_t* total, int32_t* complete) {
  if (done) *done = job->done.load();
  if (total) *total = job->total;
  if (complete) *complete = job->complete.load() ? 1 : 0;
}

void CancelJob(rad_job* job) { job->cancel.store(true); }

void WaitJob(rad_job* job, float* fb, float* sumsq, rad_job_stats* stats) {
  if (!job->joined) {
    job->th.join();
    job->joined = true;
  }
  if (job->status != RAD_OK && job->status != RAD_ERR_CANCELLED) throw StatusError(job->status, job->error);
  HipCheck(hipSetDevice(job->scene->ctx->device), "hipSetDevice");
  size_t n = size_t(job->w.npix) * 3 * sizeof(float);
  if (fb) HipCheck(hipMemcpy(fb, job->d_fb.ptr, n, hipMemcpyDeviceToHost), "copy framebuffer");
  if (sumsq) HipCheck(hipMemcpy(sumsq, job->d_sq.ptr, n, hipMemcpyDeviceToHost), "copy sum of squares");
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->render_ms = job->render_ms;
    stats->bvh_build_ms = job->scene->bvh_ms;
    stats->samples = job->stats[ST_SAMPLES];
    stats->segments = job->segments;
    stats->shadow_rays = job->shadow_rays;
    stats->medium_steps = job->stats[ST_MEDIUM];
    stats->iterations = job->iterations;
    stats->slots = job->w.Q;
    for (int k = 0; k < RAD_KERNEL_COUNT; k++) {
      stats->kernel_ms[k] = job->kernel_ms[k];
      stats->kernel_launches[k] = job->kernel_launches[k];
    }
  }
  if (job->status == RAD_ERR_CANCELLED) throw StatusError(RAD_ERR_CANCELLED, "render cancelled");
}

void DestroyJob(rad_job* job) {
  if (!job) return;
  if (!job->joined && job->th.joinable()) {
    job->cancel.store(true);
    job->th.join();
  }
  (void)hipSetDevice(job->scene->ctx->device);
  if (job->stream) (void)hipStreamDestroy(job->stream);
  if (job->h_cnt) (void)hipHostFree(job->h_cnt);
  delete job;
}

}  // namespace rad
