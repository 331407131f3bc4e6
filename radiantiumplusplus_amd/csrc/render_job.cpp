#include "render_job.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "guide_state.h"
#include "wavefront.h"

struct rad_job {
  rad_scene* scene = nullptr;
  rad_render_desc desc{};
  rad::Wavefront w{};
  rad::DevScene dev{};  // the scene's device view with this job's camera
  std::unique_ptr<rad::DeviceBuffer> arena;
  rad::DeviceBuffer d_fb, d_sq;
  rad::DeviceBuffer d_gnet;  // job-private copy of the guide weights (guiding = 2)
  std::shared_ptr<rad::guide::GuideState> guide;
  hipStream_t stream = nullptr;
  uint32_t* h_cnt = nullptr;  // pinned readback: counters + stats + guide words
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
  uint64_t segments = 0, shadow_rays = 0, suspended = 0;
  uint64_t guide_records = 0;
  float guide_loss = 0.f, guide_alpha = 0.f;
  uint64_t stats[rad::ST_WORDS] = {};
  double kernel_ms[RAD_KERNEL_COUNT] = {};
  uint32_t kernel_launches[RAD_KERNEL_COUNT] = {};
  std::vector<uint32_t> pass_rows;
  std::vector<uint8_t> pass_train;
  uint32_t* batch = nullptr;  // device minibatch indices (inside the arena)
  bool sorted = false;        // material-sorted shading
  uint32_t blocks_per_cu = 0, trace_blocks_per_cu = 0;
  uint32_t mat_mask = 0;      // BSDF types present in the scene
  uint32_t* tally = nullptr;  // RAD_RENDER_CHECK_QUEUES: tally words per extend-queue position (inside the arena)
  bool tally_selftest = false;
  bool joined = false;
};

namespace rad {

namespace {

constexpr uint64_t kDefaultSlots = 4u << 20;     // paths in flight
constexpr uint64_t kResultBudget = 3ull << 30;   // bytes of per-sample results per pass
constexpr uint32_t kRecordCap = 2u << 20;        // guide training records per pass
constexpr uint32_t kRecordSlots = 512u << 10;    // 1 in Q / kRecordSlots state indices record training data
constexpr uint32_t kTrainBatch = 8192;           // records per training step
constexpr uint32_t kTrainSteps = 96;             // training steps after each training pass
constexpr float kTrainLr = 4e-3f;
constexpr float kGuideAlpha = 0.5f;
constexpr float kTrainWmax = 64.f;               // cap of a record's normalised weight
constexpr uint32_t kFirstTrainPaths = 2u << 20;   // paths in the first training pass
constexpr float kGuideMinRough = 0.f;            // glossy BSDFs below this alpha are not guided
constexpr uint32_t kGuideMaxDepth = 2;          // deepest guided vertex (bounces before it)
constexpr uint32_t kVisitBudget = 1024;          // node visits of a path ray per extend launch
constexpr uint32_t kResumeCap = 16384;           // suspended traversals per extend launch
constexpr int kIterBatch = 4;                    // wavefront iterations enqueued per host wait
// pinned readback words: counters + stats + guide words, then one counter and
// stats area per iteration of a batch
constexpr uint32_t kIterWords = CNT_WORDS + 2 * ST_WORDS;
constexpr uint32_t kIterReadback = CNT_WORDS + 2 * ST_WORDS + 2 * G_WORDS;

// Training hyper-parameters; RAD_GUIDE_STEPS / RAD_GUIDE_LR / RAD_GUIDE_ALPHA
// in the environment override the defaults above.
uint32_t TrainSteps() {
  const char* e = std::getenv("RAD_GUIDE_STEPS");
  return e ? uint32_t(std::max(1, std::atoi(e))) : kTrainSteps;
}
float TrainLr() {
  const char* e = std::getenv("RAD_GUIDE_LR");
  return e ? float(std::atof(e)) : kTrainLr;
}
float TrainWmax() {
  const char* e = std::getenv("RAD_GUIDE_WMAX");
  return e ? float(std::atof(e)) : kTrainWmax;
}
float GuideMinRough() {
  const char* e = std::getenv("RAD_GUIDE_MINROUGH");
  return e ? float(std::atof(e)) : kGuideMinRough;
}
// Deepest guided vertex (bounces before it), RAD_GUIDE_DEPTH (-1: all).
// Default 2: on the door-gap room guiding the first three vertices gives the
// error of guiding every vertex at less inference and shading (time to the
// plain tracer's relMSE 0.70-0.74x plain against 0.77x for every vertex;
// depth 3: 0.76x; depth 1: 0.85x; depth 0: 0.92x).
uint32_t GuideMaxDepth() {
  const char* e = std::getenv("RAD_GUIDE_DEPTH");
  if (!e) return kGuideMaxDepth;
  const int v = std::atoi(e);
  return v < 0 ? 0xffffffffu : uint32_t(v);
}
// Share of the samples spent in training passes, 1 / RAD_GUIDE_TRAIN_DIV
// (default 8: on the door-gap room 1/4 gave the same error at 2 % more time,
// 1/16 a higher error).
uint32_t TrainDiv() {
  const char* e = std::getenv("RAD_GUIDE_TRAIN_DIV");
  return e ? uint32_t(std::max(1, std::atoi(e))) : 8u;
}
// RAD_VISIT_BUDGET overrides kVisitBudget; 0 lets every traversal finish in
// the launch that started it.
uint32_t VisitBudget() {
  const char* e = std::getenv("RAD_VISIT_BUDGET");
  return e ? uint32_t(std::max(0, std::atoi(e))) : kVisitBudget;
}
float GuideAlpha() {
  const char* e = std::getenv("RAD_GUIDE_ALPHA");
  return e ? std::min(0.95f, std::max(0.05f, float(std::atof(e)))) : kGuideAlpha;
}
size_t Align(size_t v) { return (v + 255) & ~size_t(255); }

struct JobShape {
  uint32_t npix, rows, Q;
  bool guided;
  uint32_t resume_cap;  // 0: no suspended traversals (vol_path, or budget 0)
  uint64_t bytes;
};

void ImageSize(const rad_scene* scene, const rad_render_desc& desc, int32_t& W, int32_t& H) {
  if (desc.width < 0 || desc.height < 0 || (desc.width > 0) != (desc.height > 0))
    throw StatusError(RAD_ERR_INVALID_ARG, "image size override needs both width and height");
  W = desc.width > 0 ? desc.width : scene->host.cam.width;
  H = desc.height > 0 ? desc.height : scene->host.cam.height;
}

JobShape Shape(const rad_scene* scene, const rad_render_desc& desc) {
  JobShape j{};
  int32_t W, H;
  ImageSize(scene, desc, W, H);
  j.npix = uint32_t(W) * uint32_t(H);
  const uint64_t spp = std::max(1u, desc.spp);
  uint64_t rows = kResultBudget / (uint64_t(j.npix) * 16);
  rows = std::min<uint64_t>(std::max<uint64_t>(rows, 1), spp);
  rows = std::min<uint64_t>(rows, 0x7fffffffull / j.npix);
  if (rows == 0) throw StatusError(RAD_ERR_INVALID_ARG, "image too large");
  j.rows = uint32_t(rows);
  const uint64_t pass = rows * j.npix;
  const uint32_t tuned = scene->tuned_slots.load();
  uint64_t Q = desc.path_slots > 0 ? desc.path_slots : tuned ? tuned : kDefaultSlots;
  Q = std::max<uint64_t>(1, std::min(Q, pass));
  j.Q = uint32_t(Q);
  j.guided = desc.guiding != 0;
  j.resume_cap = desc.integrator == RAD_INTEGRATOR_PATH && VisitBudget() > 0 ? uint32_t(std::min<uint64_t>(Q, kResumeCap)) : 0;
  // per path: rays and path records of the two extend queues (32 B and 64 B
  // each), hit, shadow origin / direction / contribution (16 B each); keys and
  // media of the two queues, hit geometry, dead list and material queues (4 B)
  j.bytes = Align(Q * 32) * 2 + Align(Q * 64) * 2 + Align(Q * 16) * 4 + Align(Q * 4) * (6 + kMatQueues) +
            Align(pass * 16) + Align(uint64_t(j.npix) * 16) * 2 +
            2 * Align(uint64_t(j.npix) * 12) + 4096;
  if (j.resume_cap) j.bytes += Align(Q * 4) * 2 + 2 * Align(uint64_t(j.resume_cap) * sizeof(ResumeState));
  if (desc.flags & RAD_RENDER_CHECK_QUEUES) j.bytes += Align(Q * 4) * 2;
  if (j.guided)
    j.bytes += Align(Q * 4) * 5 + Align(Q * 2 * guide::kOut) +
               Align(uint64_t(kRecordCap) * sizeof(guide::TrainRecord)) + Align(kTrainBatch * 4) + 256 +
               Align(guide::kWeights * 2) + Align(guide::kBiases * 4);
  return j;
}

// Sample passes: without training, as few passes as the result entries allow,
// of equal size; with online training, first passes of r, 2r, 4r, ... rows
// (each followed by training on the records it produced; r rows hold about
// 2M paths, enough records for a useful first step) until an eighth of the
// samples (1 / RAD_GUIDE_TRAIN_DIV) are spent, then the rest.
void PlanPasses(rad_job* job, uint32_t npix, uint32_t maxRows) {
  const uint32_t spp = job->desc.spp;
  uint32_t done = 0;
  if (job->desc.guiding == 1) {
    const uint32_t budget = std::max(1u, spp / TrainDiv());
    uint32_t r = std::max(1u, std::min(budget, (kFirstTrainPaths + npix - 1) / npix));
    while (done + r <= budget && r <= maxRows) {
      job->pass_rows.push_back(r);
      job->pass_train.push_back(1);
      done += r;
      r *= 2;
    }
  }
  const uint32_t rest = spp - done;
  const uint32_t n = (rest + maxRows - 1) / maxRows;
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t r = rest / n + (k < rest % n ? 1u : 0u);
    job->pass_rows.push_back(r);
    job->pass_train.push_back(0);
  }
}

void Train(rad_job* job, hipStream_t st) {
  Wavefront& w = job->w;
  const uint32_t* g = job->h_cnt + CNT_WORDS + 2 * ST_WORDS;
  uint32_t nrec = std::min(g[G_RECORDS], w.rec_cap);
  float wsum;
  std::memcpy(&wsum, &g[G_WSUM], 4);
  job->guide_records += nrec;
  if (nrec >= 256 && wsum > 0.f) {
    guide::GuideState& gs = *job->guide;
    const float wnorm = float(nrec) / wsum;
    const guide::GuideNet net = w.net;
    const uint32_t steps = TrainSteps();
    const float lr = TrainLr();
    const float wmax = TrainWmax();
    for (uint32_t k = 0; k < steps; k++) {
      HipCheck(hipMemsetAsync(gs.scal.ptr, 0, 4, st), "clear loss");
      guide::LaunchBatch(job->batch, kTrainBatch, nrec, gs.steps * 7919u + k + 1u, st);
      guide::LaunchTrain(net, w.recs, job->batch, kTrainBatch, wnorm, wmax, gs.grad.as<float>(), gs.scal.as<float>(), st);
      gs.Step(lr, 1.f / float(kTrainBatch), st);
    }
    HipCheck(hipGetLastError(), "guide training launch");
    float loss = 0.f;
    HipCheck(hipMemcpyAsync(&job->h_cnt[CNT_WORDS + 2 * ST_WORDS + G_WORDS], gs.scal.ptr, 4, hipMemcpyDeviceToHost, st),
             "loss readback");
    HipCheck(hipStreamSynchronize(st), "guide training");
    std::memcpy(&loss, &job->h_cnt[CNT_WORDS + 2 * ST_WORDS + G_WORDS], 4);
    job->guide_loss = loss / float(kTrainBatch);
    gs.Publish(st);
  }
  HipCheck(hipMemsetAsync(w.gcnt, 0, G_WORDS * 4, st), "clear guide counters");
}

void RunJob(rad_job* job) {
  std::unique_lock<std::mutex> trainLock;
  try {
    HipCheck(hipSetDevice(job->scene->ctx->device), "hipSetDevice");
    if (job->desc.guiding == 1) trainLock = std::unique_lock<std::mutex>(job->scene->guide_mu);
    const DevScene scene = job->dev;
    Wavefront& w = job->w;
    hipStream_t st = job->stream;
    const int cus = job->scene->ctx->cu_count;
    rad_render_desc launch = job->desc;  // resolved at submission; defaults for direct callers
    ResolveLaunch(job->scene, launch);
    const int bpc = int(launch.blocks_per_cu), tpc = int(launch.trace_blocks_per_cu);
    LaunchCfg cfg{std::max(64, cus * bpc), std::max(64, cus * tpc)};
    job->blocks_per_cu = uint32_t(bpc);
    job->trace_blocks_per_cu = uint32_t(tpc);
    const int inferBlocks = std::max(64, job->scene->ctx->cu_count * 2);  // two 512-thread blocks per CU
    const bool prof = (job->desc.flags & RAD_RENDER_PROFILE) != 0;
    const size_t readback = CNT_WORDS * sizeof(uint32_t) + ST_WORDS * sizeof(uint64_t) + G_WORDS * 4;
    hipEvent_t evs[kIterBatch][7] = {};  // stage boundaries of each iteration of a batch
    if (prof)
      for (auto& row : evs)
        for (auto& e : row) HipCheck(hipEventCreate(&e), "hipEventCreate");
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t baseFlags = w.flags & WF_QUIRKS;
    const float* lo = job->scene->host.world_lo;
    const float* hi = job->scene->host.world_hi;
    uint32_t row0 = 0;
    for (size_t pi = 0; pi < job->pass_rows.size() && job->status == RAD_OK; pi++) {
      w.row0 = row0;
      w.rows = job->pass_rows[pi];
      row0 += w.rows;
      w.flags = baseFlags;
      if (job->desc.guiding == 2) {
        w.flags |= WF_GUIDE;
        w.net.alpha = GuideAlpha();
        w.net.min_rough = GuideMinRough();
        w.net.max_depth = GuideMaxDepth();
      } else if (job->desc.guiding == 1) {
        const bool trained = job->guide->steps > 0;
        w.net = job->guide->Net(lo, hi, trained ? GuideAlpha() : 0.f, GuideMinRough());
        w.net.max_depth = GuideMaxDepth();
        if (trained) w.flags |= WF_GUIDE;
        if (job->pass_train[pi]) w.flags |= WF_RECORD;
      }
      job->guide_alpha = (w.flags & WF_GUIDE) ? w.net.alpha : 0.f;
      LaunchInitPass(w, pi == 0, st);
      int in = 0;
      bool fresh = true;  // the pass's first regen: no dead-list entry holds a finished sample
      for (bool passDone = false; !passDone;) {
        // Iterations are enqueued kIterBatch at a time and the host waits once
        // per batch: every kernel takes its work count from device counters,
        // so an iteration enqueued after the pass ran out of work does nothing.
        // Each iteration's counters go to their own readback area.
        const int batch = kIterBatch;
        int ins[kIterBatch];
        for (int b = 0; b < batch; b++) {
          ins[b] = in;
          uint32_t* const hc = job->h_cnt + kIterReadback + b * kIterWords;
          hipEvent_t* const ev = evs[b];
          if (prof) HipCheck(hipEventRecord(ev[0], st), "event");
          LaunchRegen(w, scene, in, fresh, cfg, st);
          fresh = false;
          LaunchReset(w, 1 - in, st);
          if (prof) HipCheck(hipEventRecord(ev[1], st), "event");
          LaunchExtend(w, scene, in, cfg, st);
          if (prof) HipCheck(hipEventRecord(ev[2], st), "event");
          const bool vol = w.integrator == RAD_INTEGRATOR_VOL_PATH;
          const bool guided = (w.flags & WF_GUIDE) != 0;
          if (job->sorted) {
            LaunchClassify(w, scene, in, cfg, st);  // with guiding or records, also the guided queues
            if (job->tally) LaunchQueueCheck(w, scene, in, job->tally, job->tally_selftest, cfg, st);
          }
          else if (!vol && guided)
            LaunchGuidePrep(w, scene, in, cfg, st);
          if (prof) HipCheck(hipEventRecord(ev[3], st), "event");
          if (!vol && guided) LaunchGuideInfer(w, scene, in, inferBlocks, job->sorted, st);
          if (prof) HipCheck(hipEventRecord(ev[4], st), "event");
          if (vol && job->sorted)
            LaunchShadeVolMaterials(w, scene, in, job->mat_mask, cfg, st);
          else if (vol)
            LaunchShadeVol(w, scene, in, cfg, st);
          else if (job->sorted)
            LaunchShadeMaterials(w, scene, in, job->mat_mask, cfg, st);
          else
            LaunchShadePath(w, scene, in, cfg, st);
          if (prof) HipCheck(hipEventRecord(ev[5], st), "event");
          if (vol)
            LaunchVolShadow(w, scene, cfg, st);
          else
            LaunchShadow(w, scene, cfg, st);
          if (prof) HipCheck(hipEventRecord(ev[6], st), "event");
          HipCheck(hipGetLastError(), "wavefront launch");
          HipCheck(hipMemcpyAsync(hc, w.cnt, CNT_WORDS * sizeof(uint32_t) + ST_WORDS * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, st),
                   "counter readback");
          if (w.gcnt)
            HipCheck(hipMemcpyAsync(job->h_cnt + CNT_WORDS + 2 * ST_WORDS, w.gcnt, G_WORDS * 4, hipMemcpyDeviceToHost, st),
                     "guide counter readback");
          in = 1 - in;
        }
        HipCheck(hipStreamSynchronize(st), "wavefront iterations");
        for (int b = 0; b < batch && !passDone; b++) {
          const uint32_t* const hc = job->h_cnt + kIterReadback + b * kIterWords;
          const int in = ins[b];
          const bool guided = (w.flags & WF_GUIDE) != 0;
          const hipEvent_t* const ev = evs[b];
          job->iterations++;
          if (prof) {
            // event intervals -> stages; classify (or guide-queue gathering) counts as shading
            const int slot[6] = {RAD_KERNEL_REGEN, RAD_KERNEL_EXTEND, RAD_KERNEL_SHADE, RAD_KERNEL_GUIDE, RAD_KERNEL_SHADE,
                                 RAD_KERNEL_SHADOW};
            for (int k = 0; k < 6; k++) {
              if (slot[k] == RAD_KERNEL_GUIDE && !guided) continue;
              float ms = 0.f;
              HipCheck(hipEventElapsedTime(&ms, ev[k], ev[k + 1]), "event time");
              job->kernel_ms[slot[k]] += ms;
              if (k != 2) job->kernel_launches[slot[k]]++;
            }
          }
          const uint64_t* hs = reinterpret_cast<const uint64_t*>(hc + CNT_WORDS);
          job->done.store(hs[ST_SAMPLES]);
          // a suspended traversal counts once, in the launch that finishes it
          job->segments += hc[in] - std::min(hc[CNT_SUSP + 1 - in], w.resume_cap);
          job->suspended += std::min(hc[CNT_SUSP + 1 - in], w.resume_cap);
          job->shadow_rays += hc[CNT_SHADOW];
          const uint32_t cont = hc[1 - in], dead = hc[CNT_DEAD];
          if (cont == 0 && dead == 0) passDone = true;
        }
        if (job->iterations > 1000000u)
          throw StatusError(RAD_ERR_STATE, "path lengths exceed one million bounces (no termination criterion)");
        if (!passDone && job->cancel.load()) {
          job->status = RAD_ERR_CANCELLED;
          break;
        }
      }
      if (job->status != RAD_OK) break;
      LaunchAccum(w, st);
      job->passes++;
      if (w.flags & WF_RECORD) {
        auto tt = std::chrono::steady_clock::now();
        Train(job, st);
        if (prof) {
          job->kernel_ms[RAD_KERNEL_TRAIN] +=
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tt).count();
          job->kernel_launches[RAD_KERNEL_TRAIN]++;
        }
      }
    }
    LaunchFinal(w, job->d_fb.as<float>(), job->d_sq.as<float>(), st);
    HipCheck(hipGetLastError(), "resolve launch");
    HipCheck(hipStreamSynchronize(st), "resolve");
    job->render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    HipCheck(hipMemcpy(job->h_cnt, w.cnt, CNT_WORDS * sizeof(uint32_t) + ST_WORDS * sizeof(uint64_t),
                       hipMemcpyDeviceToHost),
             "stats readback");
    (void)readback;
    const uint64_t* hs = reinterpret_cast<const uint64_t*>(job->h_cnt + CNT_WORDS);
    for (int i = 0; i < ST_WORDS; i++) job->stats[i] = hs[i];
    if (prof)
      for (auto& row : evs)
        for (auto& e : row) (void)hipEventDestroy(e);
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

void ResolveLaunch(const rad_scene* scene, rad_render_desc& d) {
  if (d.path_slots == 0) {
    const uint32_t t = scene->tuned_slots.load();
    d.path_slots = t ? t : uint32_t(kDefaultSlots);
  }
  if (d.blocks_per_cu == 0) {
    const uint32_t t = scene->tuned_blocks_per_cu.load();
    d.blocks_per_cu = t ? t : 8u;
  }
  if (d.trace_blocks_per_cu == 0) {
    const uint32_t t = scene->tuned_trace_blocks_per_cu.load();
    d.trace_blocks_per_cu = t ? t : 8u;
  }
}

uint64_t JobFootprint(const rad_scene* scene, const rad_render_desc& desc, uint32_t* slots) {
  JobShape j = Shape(scene, desc);
  if (slots) *slots = j.Q;
  return j.bytes;
}

rad_job* StartJob(rad_scene* scene, const rad_render_desc& desc, std::unique_ptr<DeviceBuffer> reuse) {
  if (desc.spp == 0) throw StatusError(RAD_ERR_INVALID_ARG, "spp must be positive");
  if (desc.integrator != RAD_INTEGRATOR_PATH && desc.integrator != RAD_INTEGRATOR_VOL_PATH)
    throw StatusError(RAD_ERR_INVALID_ARG, "unknown integrator");
  if (desc.guiding < 0 || desc.guiding > 2) throw StatusError(RAD_ERR_INVALID_ARG, "guiding must be 0, 1 or 2");
  if (desc.guiding && desc.integrator != RAD_INTEGRATOR_PATH)
    throw StatusError(RAD_ERR_UNSUPPORTED, "learned guiding is available for the path integrator");
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
  job->dev = scene->dev;
  {
    int32_t W, H;
    ImageSize(scene, desc, W, H);
    if (W != scene->host.cam.width || H != scene->host.cam.height) {
      rad_camera_desc c = scene->host.camera_desc;
      c.width = W;
      c.height = H;
      job->dev.cam = MakeCameraRec(c);
    }
  }
  const JobShape js = Shape(scene, desc);
  const uint32_t npix = js.npix, Q = js.Q;
  job->total = uint64_t(npix) * desc.spp;
  PlanPasses(job.get(), js.npix, js.rows);
  // one arena for every per-path, per-sample and per-pixel array
  if (reuse && reuse->bytes >= js.bytes) {
    job->arena = std::move(reuse);
  } else {
    job->arena = std::make_unique<DeviceBuffer>();
    job->arena->alloc(js.bytes);
  }
  char* p = static_cast<char*>(job->arena->ptr);
  char* const end = p + job->arena->bytes;
  auto take = [&](size_t n) {
    char* r = p;
    p += Align(n);
    if (p > end) throw StatusError(RAD_ERR_HIP, "render arena layout exceeds its size");
    return r;
  };
  Wavefront& w = job->w;
  const size_t f4 = size_t(Q) * 16, u4 = size_t(Q) * 4;
  w.ray[0] = (float4*)take(2 * f4);
  w.ray[1] = (float4*)take(2 * f4);
  float4* rec = (float4*)take(8 * f4);  // 2 Q records
  w.thr.p = rec;
  w.lsum.p = rec + 1;
  w.prev.p = rec + 2;
  w.meta.p = reinterpret_cast<uint4*>(rec + 3);
  w.hit = (float4*)take(f4);
  w.sh_o = (float4*)take(f4);
  w.sh_d = (float4*)take(f4);
  w.sh_c = (float4*)take(f4);
  w.key = (uint32_t*)take(2 * u4);
  w.medium = (uint32_t*)take(2 * u4);
  w.hit_geom = (uint32_t*)take(u4);
  w.dead = (uint32_t*)take(u4);
  for (int k = 0; k < kMatQueues; k++) w.mq[k] = (uint32_t*)take(u4);
  w.res = (float4*)take(size_t(js.rows) * npix * 16);
  w.pacc = (float4*)take(size_t(npix) * 16);
  w.pacc2 = (float4*)take(size_t(npix) * 16);
  w.cnt = (uint32_t*)take(CNT_WORDS * 4 + ST_WORDS * 8);
  w.stats = reinterpret_cast<unsigned long long*>(w.cnt + CNT_WORDS);
  if (js.resume_cap) {
    w.susp = (uint32_t*)take(2 * u4);
    w.resume[0] = (ResumeState*)take(size_t(js.resume_cap) * sizeof(ResumeState));
    w.resume[1] = (ResumeState*)take(size_t(js.resume_cap) * sizeof(ResumeState));
    w.resume_cap = js.resume_cap;
    w.visit_budget = VisitBudget();
  }
  if (desc.flags & RAD_RENDER_CHECK_QUEUES) {
    job->tally = (uint32_t*)take(u4);
    w.passsrc = (uint32_t*)take(u4);
    job->tally_selftest = (desc.flags & RAD_RENDER_CHECK_SELFTEST) != 0;
    HipCheck(hipMemset(job->tally, 0, u4), "queue check tally");
  }
  if (js.guided) {
    w.gq = (uint32_t*)take(u4);
    w.gpend = (uint32_t*)take(2 * u4);
    w.glast = (uint32_t*)take(2 * u4);
    w.graw = (_Float16*)take(size_t(Q) * 2 * guide::kOut);
    w.recs = (guide::TrainRecord*)take(size_t(kRecordCap) * sizeof(guide::TrainRecord));
    job->batch = (uint32_t*)take(kTrainBatch * 4);
    w.gcnt = (uint32_t*)take(256);
    w.rec_cap = kRecordCap;
    w.rec_stride = std::max(1u, Q / kRecordSlots);
    _Float16* wh = (_Float16*)take(guide::kWeights * 2);
    float* bias = (float*)take(guide::kBiases * 4);
    {
      std::lock_guard<std::mutex> lk(scene->mu);
      if (!scene->guide) scene->guide = std::make_shared<guide::GuideState>();
      job->guide = scene->guide;
    }
    if (desc.guiding == 2) {  // snapshot of the published weights: concurrent training cannot touch this render
      job->guide->Snapshot(wh, bias);
      w.net = job->guide->Net(scene->host.world_lo, scene->host.world_hi, GuideAlpha(), GuideMinRough());
      w.net.w = wh;
      w.net.b = bias;
    }
    HipCheck(hipMemset(w.gcnt, 0, 256), "guide counters");
  }
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
  job->sorted = !(desc.flags & RAD_RENDER_UNSORTED);
  for (const BsdfRec& b : scene->host.bsdfs) job->mat_mask |= 1u << (b.tex_any ? kMatOther : std::min<int32_t>(b.type, kMatOther));
  HipCheck(hipHostMalloc((void**)&job->h_cnt, (kIterReadback + kIterBatch * kIterWords) * 4, hipHostMallocDefault),
           "hipHostMalloc");
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
    stats->suspended = job->suspended;
    stats->queue_checked = job->stats[ST_QCHECK];
    stats->queue_errors = job->stats[ST_QERR];
    stats->shadow_rays = job->shadow_rays;
    stats->medium_steps = job->stats[ST_MEDIUM];
    stats->iterations = job->iterations;
    stats->slots = job->w.Q;
    for (int k = 0; k < RAD_KERNEL_COUNT; k++) {
      stats->kernel_ms[k] = job->kernel_ms[k];
      stats->kernel_launches[k] = job->kernel_launches[k];
    }
    stats->passes = job->passes;
    stats->guide_steps = job->guide ? job->guide->steps : 0;
    stats->guide_records = job->guide_records;
    stats->guide_loss = job->guide_loss;
    stats->guide_alpha = job->guide_alpha;
    stats->blocks_per_cu = job->blocks_per_cu;
    stats->trace_blocks_per_cu = job->trace_blocks_per_cu;
  }
  if (job->status == RAD_ERR_CANCELLED) throw StatusError(RAD_ERR_CANCELLED, "render cancelled");
}

std::unique_ptr<DeviceBuffer> TakeArena(rad_job* job) {
  if (!job->joined && job->th.joinable()) {
    job->cancel.store(true);
    job->th.join();
    job->joined = true;
  }
  return std::move(job->arena);
}

bool JobComplete(const rad_job* job) { return job->complete.load(); }

void DestroyJob(rad_job* job) {
  if (!job) return;
  if (!job->joined && job->th.joinable()) {
    job->cancel.store(true);
    job->th.join();
  }
  (void)hipSetDevice(job->scene->ctx->device);
  if (job->stream) {
    (void)hipStreamSynchronize(job->stream);  // nothing of this job may still write the arena
    (void)hipStreamDestroy(job->stream);
  }
  if (job->h_cnt) (void)hipHostFree(job->h_cnt);
  if (job->arena) {  // keep the larger arena on the scene for its next job
    rad_scene* sc = job->scene;
    std::lock_guard<std::mutex> lk(sc->arena_mu);
    if (!sc->spare_arena || sc->spare_arena->bytes < job->arena->bytes) sc->spare_arena = std::move(job->arena);
  }
  delete job;
}

std::unique_ptr<DeviceBuffer> TakeSpareArena(rad_scene* scene) {
  std::lock_guard<std::mutex> lk(scene->arena_mu);
  return std::move(scene->spare_arena);
}

}  // namespace rad
