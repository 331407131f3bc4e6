// Render-serving loop.  Requests (a resident scene + a render description)
// arrive at any time; a scheduler thread admits them in arrival order while
// their HBM footprint fits the budget, starts each as a render job on its own
// stream (so the kernels of admitted requests overlap on the device), and
// retires finished jobs, whose arenas go back to a pool that later requests
// reuse.  Scenes (BVHs, vertex data, density grids, guide weights) stay
// resident between requests.  Each request keeps submit / admit / finish
// times; the server reports p50 / p99 latency beside throughput.
//
// A request's image depends only on its scene and description (the RNG key
// includes request_id; every sample has its own result entry), so the same
// request renders the same image alone or in any mix of requests.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "render_job.h"

namespace rad {
void SetLastError(const std::string& msg);
}

namespace {

using Clock = std::chrono::steady_clock;

double Ms(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

struct Request {
  uint64_t id = 0;
  rad_scene* scene = nullptr;
  rad_render_desc desc{};
  uint64_t footprint = 0;
  uint64_t held = 0;  // arena bytes this request holds in the server's account
  uint32_t width = 0, height = 0;
  int32_t state = RAD_REQ_QUEUED;
  rad_status status = RAD_OK;
  std::string error;
  Clock::time_point t_submit, t_admit, t_done;
  rad_job* job = nullptr;
  std::vector<float> fb, sq;
  rad_job_stats stats{};
};

}  // namespace

struct rad_server {
  rad_ctx* ctx = nullptr;
  uint64_t budget = 0;
  uint32_t max_inflight = 0;
  std::mutex mu;
  std::condition_variable cv_done;  // a request finished
  std::condition_variable cv_work;  // a request arrived or the server stops
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::deque<uint64_t> queue;
  std::vector<uint64_t> running;
  std::vector<std::unique_ptr<rad::DeviceBuffer>> pool;  // arenas of retired jobs
  uint64_t held_bytes = 0;  // arenas in use + pooled
  uint64_t peak_held = 0;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread th;
  // statistics
  uint64_t submitted = 0, completed = 0, failed = 0, samples = 0;
  std::vector<double> latencies;
  bool any = false;
  Clock::time_point t_first, t_last;
};

namespace {

using namespace rad;

// Take a pooled arena that fits, or make room for a new one; false = wait.
// `held` = the bytes the request will hold (the pooled arena's size, or
// `need` for a fresh allocation).
bool Reserve(rad_server* s, uint64_t need, std::unique_ptr<DeviceBuffer>& out, uint64_t& held) {
  int best = -1;
  for (size_t i = 0; i < s->pool.size(); i++)
    if (s->pool[i]->bytes >= need && (best < 0 || s->pool[i]->bytes < s->pool[best]->bytes)) best = int(i);
  if (best >= 0) {
    out = std::move(s->pool[best]);
    s->pool.erase(s->pool.begin() + best);
    held = out->bytes;  // already in held_bytes as a pooled arena
    return true;
  }
  // free pooled arenas (largest first) until the new one fits the budget
  std::sort(s->pool.begin(), s->pool.end(), [](const auto& a, const auto& b) { return a->bytes < b->bytes; });
  while (s->held_bytes + need > s->budget && !s->pool.empty()) {
    s->held_bytes -= s->pool.back()->bytes;
    s->pool.pop_back();
  }
  if (s->held_bytes + need > s->budget && !s->running.empty()) return false;
  s->held_bytes += need;  // the job allocates it
  s->peak_held = std::max(s->peak_held, s->held_bytes);
  held = need;
  out.reset();
  return true;
}

void Retire(rad_server* s, Request& r) {
  r.fb.assign(size_t(r.width) * r.height * 3, 0.f);
  if (r.desc.flags & RAD_RENDER_WANT_SUMSQ) r.sq.assign(r.fb.size(), 0.f);
  try {
    WaitJob(r.job, r.fb.data(), r.sq.empty() ? nullptr : r.sq.data(), &r.stats);
    r.state = RAD_REQ_DONE;
  } catch (const StatusError& e) {
    r.status = e.code;
    r.error = e.what();
    r.state = RAD_REQ_FAILED;
  } catch (const std::exception& e) {
    r.status = RAD_ERR_HIP;
    r.error = e.what();
    r.state = RAD_REQ_FAILED;
  }
  std::unique_ptr<DeviceBuffer> a = TakeArena(r.job);
  DestroyJob(r.job);
  r.job = nullptr;
  if (a && a->bytes == r.held)
    s->pool.push_back(std::move(a));  // stays in held_bytes, now pooled
  else
    s->held_bytes -= std::min(s->held_bytes, r.held);
  r.t_done = Clock::now();
  if (r.state == RAD_REQ_DONE) {
    s->completed++;
    s->samples += r.stats.samples;
    s->latencies.push_back(Ms(r.t_submit, r.t_done));
  } else {
    s->failed++;
  }
  s->t_last = r.t_done;
}

void Loop(rad_server* s) {
  (void)hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    // retire finished jobs
    for (size_t i = 0; i < s->running.size();) {
      Request& r = *s->reqs[s->running[i]];
      if (JobComplete(r.job)) {
        Retire(s, r);
        s->running.erase(s->running.begin() + i);
        s->cv_done.notify_all();
      } else {
        i++;
      }
    }
    // admit in arrival order while the footprint fits
    while (!s->queue.empty() && s->running.size() < s->max_inflight) {
      Request& r = *s->reqs[s->queue.front()];
      std::unique_ptr<DeviceBuffer> arena;
      if (!Reserve(s, r.footprint, arena, r.held)) break;
      s->queue.pop_front();
      r.t_admit = Clock::now();
      try {
        r.job = StartJob(r.scene, r.desc, std::move(arena));
        r.state = RAD_REQ_RUNNING;
        s->running.push_back(r.id);
      } catch (const StatusError& e) {
        r.status = e.code;
        r.error = e.what();
        r.state = RAD_REQ_FAILED;
        s->failed++;
        s->held_bytes -= std::min(s->held_bytes, r.held);
        s->cv_done.notify_all();
      } catch (const std::exception& e) {
        r.status = RAD_ERR_HIP;
        r.error = e.what();
        r.state = RAD_REQ_FAILED;
        s->failed++;
        s->held_bytes -= std::min(s->held_bytes, r.held);
        s->cv_done.notify_all();
      }
    }
    if (s->running.empty() && s->queue.empty())
      s->cv_work.wait(lk);
    else
      s->cv_work.wait_for(lk, std::chrono::microseconds(200));
  }
}

double Quantile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  double pos = q * double(v.size() - 1);
  size_t i = size_t(pos);
  double f = pos - double(i);
  return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

rad_status Fail(rad_status st, const std::string& m) {
  SetLastError(m);
  return st;
}

}  // namespace

extern "C" {

rad_status rad_hip_server_create(rad_ctx* ctx, const rad_server_opts* opts, rad_server** out) {
  if (!ctx || !out) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  auto s = std::make_unique<rad_server>();
  s->ctx = ctx;
  s->budget = opts && opts->hbm_budget ? opts->hbm_budget : (uint64_t(32) << 30);
  s->max_inflight = opts && opts->max_inflight ? opts->max_inflight : 8;
  s->th = std::thread(Loop, s.get());
  *out = s.release();
  return RAD_OK;
}

rad_status rad_hip_server_submit(rad_server* s, rad_scene* scene, const rad_render_desc* desc, uint64_t* request) {
  if (!s || !scene || !desc || !request) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  if (!scene->bvh_ready) return Fail(RAD_ERR_STATE, "scene has no BVH (rad_hip_bvh_build)");
  auto r = std::make_unique<Request>();
  r->desc = *desc;
  ResolveLaunch(scene, r->desc);
  try {
    r->footprint = JobFootprint(scene, r->desc, nullptr);
  } catch (const StatusError& e) {
    return Fail(e.code, e.what());
  }
  if (r->footprint > s->budget) return Fail(RAD_ERR_OUT_OF_MEMORY, "request footprint exceeds the server's HBM budget");
  r->scene = scene;
  r->width = uint32_t(desc->width > 0 ? desc->width : scene->host.cam.width);
  r->height = uint32_t(desc->height > 0 ? desc->height : scene->host.cam.height);
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  if (r->desc.request_id == 0) r->desc.request_id = uint32_t(r->id);
  r->t_submit = Clock::now();
  if (!s->any) {
    s->any = true;
    s->t_first = r->t_submit;
  }
  *request = r->id;
  s->queue.push_back(r->id);
  s->reqs[r->id] = std::move(r);
  s->submitted++;
  s->cv_work.notify_all();
  return RAD_OK;
}

rad_status rad_hip_server_poll(rad_server* s, uint64_t request, int32_t* state) {
  if (!s || !state) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(request);
  if (it == s->reqs.end()) return Fail(RAD_ERR_INVALID_ARG, "unknown request");
  *state = it->second->state;
  return RAD_OK;
}

rad_status rad_hip_server_wait(rad_server* s, uint64_t request, float* fb, float* sumsq, rad_job_stats* stats,
                               rad_request_times* times) {
  if (!s) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(request);
  if (it == s->reqs.end()) return Fail(RAD_ERR_INVALID_ARG, "unknown request");
  Request* r = it->second.get();
  s->cv_done.wait(lk, [&] { return r->state == RAD_REQ_DONE || r->state == RAD_REQ_FAILED; });
  if (times) {
    times->queue_ms = Ms(r->t_submit, r->t_admit);
    times->run_ms = Ms(r->t_admit, r->t_done);
    times->latency_ms = Ms(r->t_submit, r->t_done);
  }
  rad_status st = r->status;
  std::string err = r->error;
  if (st == RAD_OK) {
    if (fb) std::copy(r->fb.begin(), r->fb.end(), fb);
    if (sumsq && !r->sq.empty()) std::copy(r->sq.begin(), r->sq.end(), sumsq);
    if (stats) *stats = r->stats;
  }
  s->reqs.erase(it);
  if (st != RAD_OK) return Fail(st, err);
  return RAD_OK;
}

rad_status rad_hip_server_stats(rad_server* s, rad_server_stats* out) {
  if (!s || !out) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = rad_server_stats{};
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->failed = s->failed;
  out->queued = s->queue.size();
  out->running = s->running.size();
  out->samples = s->samples;
  out->wall_ms = s->any && s->completed ? Ms(s->t_first, s->t_last) : 0.0;
  out->samples_per_s = out->wall_ms > 0 ? double(s->samples) / (out->wall_ms * 1e-3) : 0.0;
  out->p50_ms = Quantile(s->latencies, 0.50);
  out->p99_ms = Quantile(s->latencies, 0.99);
  double sum = 0;
  for (double v : s->latencies) sum += v;
  out->mean_ms = s->latencies.empty() ? 0.0 : sum / double(s->latencies.size());
  out->held_bytes = s->held_bytes;
  out->peak_held_bytes = s->peak_held;
  out->budget_bytes = s->budget;
  return RAD_OK;
}

void rad_hip_server_destroy(rad_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  s->th.join();
  for (auto& kv : s->reqs) {
    if (kv.second->job) {
      CancelJob(kv.second->job);
      DestroyJob(kv.second->job);
    }
  }
  delete s;
}

}  // extern "C"
