// Asynchronous render jobs: one host thread and one HIP stream per job drive
// the wavefront loop (the GPU analogue of SampleRenderer::Start's render
// thread, sample_renderer.cpp:60-122).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/rad_hip.h"
#include "scene_host.h"

namespace rad {

struct StatusError : std::runtime_error {
  rad_status code;
  StatusError(rad_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// `reuse`: an arena from an earlier job, taken when large enough.
// Fill path_slots / blocks_per_cu / trace_blocks_per_cu left 0 with the
// scene's tuned values (or the defaults), so that a request's footprint and
// its job see one launch configuration even if the autotuner runs between.
void ResolveLaunch(const rad_scene* scene, rad_render_desc& desc);
rad_job* StartJob(rad_scene* scene, const rad_render_desc& desc, std::unique_ptr<DeviceBuffer> reuse = nullptr);
bool JobComplete(const rad_job* job);
// Detach the job's arena (joins its thread first) so a later job can reuse it.
std::unique_ptr<DeviceBuffer> TakeArena(rad_job* job);
// The scene's spare arena (left by the last destroyed job), or null.
std::unique_ptr<DeviceBuffer> TakeSpareArena(rad_scene* scene);
void PollJob(rad_job* job, uint64_t* done, uint64_t* total, int32_t* complete);
void CancelJob(rad_job* job);
void WaitJob(rad_job* job, float* fb, float* sumsq, rad_job_stats* stats);
void DestroyJob(rad_job* job);

// Device bytes a job of this shape needs (admission control).
uint64_t JobFootprint(const rad_scene* scene, const rad_render_desc& desc, uint32_t* slots_per_pixel);

void UploadScene(rad_scene* s);

}  // namespace rad
