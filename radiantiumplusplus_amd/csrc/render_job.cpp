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
  uint64_t stats[rad::ST_WORDS] = {};
  double kernel_ms[RAD_KERNEL_COUNT] = {};
  uint32_t kernel_launches[RAD_KERNEL_COUNT] = {};
  bool joined = false;
};

namespace rad {

namespace {

uint32_t ChooseSlotsPerPixel(uint32_t npix, uint32_t spp, uint32_t requested) {
  if (requested > 0) return std::max(1u, std::min(requested, spp));
  const uint64_t target = 4u << 20;  // ~4M paths in flight
  uint32_t best = 1;
  for (uint32_t k = 1; k <= spp; k++) {
    if (spp % k) continue;
    if (uint64_t(npix) * k <= std::max<uint64_t>(target, npix)) best = k;
  }
  return best;
}

constexpr size_t kSlotBytes = 16 * 12 + 4 * 6;  // float4 arrays + uint32 arrays per slot

size_t Align(size_t v) { return (v + 255) & ~size_t(255); }

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
    LaunchInit(w, st);
    int in = 0;
    for (;;) {
      if (prof) HipCheck(hipEventRecord(ev[0], st), "event");
      LaunchRegen(w, scene, in, cfg, st);
      LaunchReset(w, 1 - in, st);
      if (prof) HipCheck(hipEventRecord(ev[1], st), "event");
      LaunchExtend(w, scene, in, cfg, st);
      if (prof) HipCheck(hipEventRecord(ev[2], st), "event");
      if (w.integrator == RAD_INTEGRATOR_PATH) {
        LaunchShadePath(w, scene, in, cfg, st);
      } else {
        throw StatusError(RAD_ERR_UNSUPPORTED, "integrator not available in this build");
      }
      if (prof) HipCheck(hipEventRecord(ev[3], st), "event");
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
    LaunchResolve(w, job->d_fb.as<float>(), job->d_sq.as<float>(), st);
    HipCheck(hipGetLastError(), "resolve launch");
    HipCheck(hipStreamSynchronize(st), "resolve");
    job->render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
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

uint64_t JobFootprint(const rad_scene* scene, const rad_render_desc& desc, uint32_t* kOut) {
  uint32_t npix = uint32_t(scene->host.cam.width) * uint32_t(scene->host.cam.height);
  uint32_t K = ChooseSlotsPerPixel(npix, std::max(1u, desc.spp), desc.slots_per_pixel);
  if (kOut) *kOut = K;
  uint64_t Q = uint64_t(npix) * K;
  return Q * kSlotBytes + 2ull * npix * 3 * sizeof(float) + 64 * 1024;
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
  auto job = std::make_unique<rad_job>();
  job->scene = scene;
  job->desc = desc;
  uint32_t npix = uint32_t(scene->host.cam.width) * uint32_t(scene->host.cam.height);
  uint32_t K = 1;
  JobFootprint(scene, desc, &K);
  uint64_t Q64 = uint64_t(npix) * K;
  if (Q64 > 0x7fffffffull) throw StatusError(RAD_ERR_INVALID_ARG, "too many path slots");
  uint32_t Q = uint32_t(Q64);
  job->total = uint64_t(npix) * desc.spp;
  // one arena for every per-slot array
  size_t f4 = Align(size_t(Q) * 16), u4 = Align(size_t(Q) * 4);
  size_t bytes = 12 * f4 + 6 * u4 + Align(CNT_WORDS * 4 + ST_WORDS * 8);
  job->arena.alloc(bytes);
  char* p = static_cast<char*>(job->arena.ptr);
  auto take = [&](size_t n) {
    char* r = p;
    p += n;
    return r;
  };
  Wavefront& w = job->w;
  w.ray_o = (float4*)take(f4);
  w.ray_d = (float4*)take(f4);
  w.thr = (float4*)take(f4);
  w.lsum = (float4*)take(f4);
  w.prev = (float4*)take(f4);
  w.meta = (uint4*)take(f4);
  w.acc = (float4*)take(f4);
  w.acc2 = (float4*)take(f4);
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
  w.cnt = (uint32_t*)take(Align(CNT_WORDS * 4 + ST_WORDS * 8));
  w.stats = reinterpret_cast<unsigned long long*>(w.cnt + CNT_WORDS);
  job->d_sq.alloc(size_t(npix) * 3 * sizeof(float));
  job->d_fb.alloc(size_t(npix) * 3 * sizeof(float));
  w.npix = npix;
  w.K = K;
  w.Q = Q;
  w.spp = desc.spp;
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
    stats->segments = job->stats[ST_SEGMENTS];
    stats->shadow_rays = job->stats[ST_SHADOW];
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
