// C++ host interface over the C ABI, shaped like the reference's renderer
// plugin API so a reference-side caller can switch to the GPU path:
//   Renderer       ~ Rad::Renderer (render/renderer.h:19-48): Start (async),
//                    Wait, Stop, IsComplete, AllTaskCount, CompletedTaskCount,
//                    ElapsedTime (ms), SaveResult (<name>.exr)
//   RendererFactory / FactoryManager ~ build/factory.h:84-89 and
//                    core/factory.cpp:5-19 (duplicate names and unknown
//                    names throw)
// The factories "path" and "vol_path" build GPU renderers for the scenes of
// the reference's JSON format (renderer.type); renderer.device "hip" is the
// default and the only device this library provides.
#pragma once

#include <chrono>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rad_hip.h"

namespace rad {
namespace offline {

// Non-zero rad_status from the C ABI, with the library's message
// (the reference's RadException family, core/types.h:59-116).
struct RadException : std::runtime_error {
  rad_status status;
  RadException(rad_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};
void Check(rad_status s, const char* what);

// A scene loaded from JSON and resident on the device.
class DeviceScene {
 public:
  DeviceScene(rad_ctx* ctx, const std::string& jsonPath);
  ~DeviceScene();
  DeviceScene(const DeviceScene&) = delete;
  DeviceScene& operator=(const DeviceScene&) = delete;
  rad_scene* handle() const { return scene_; }
  const rad_render_desc& render_desc() const { return *render_; }
  std::string renderer_type() const;
  std::string device() const;
  int width() const;
  int height() const;
  double bvh_build_ms() const;

 private:
  rad_loaded_scene* loaded_ = nullptr;
  rad_scene* scene_ = nullptr;
  const rad_render_desc* render_ = nullptr;
};

class Renderer {
 public:
  Renderer(std::shared_ptr<DeviceScene> scene, const rad_render_desc& desc);
  virtual ~Renderer();
  Renderer(const Renderer&) = delete;
  Renderer& operator=(const Renderer&) = delete;

  void Start();
  void Wait();
  void Stop();
  bool IsComplete() const;
  uint64_t AllTaskCount() const;
  uint64_t CompletedTaskCount() const;
  int64_t ElapsedTime() const;  // milliseconds since Start
  const DeviceScene& GetScene() const { return *scene_; }
  // Writes <directory>/<name>.exr (sample_renderer.cpp:34-46).
  void SaveResult(const std::string& directory, const std::string& name) const;
  const std::vector<float>& Image() const { return image_; }
  const rad_job_stats& Stats() const { return stats_; }

 protected:
  std::shared_ptr<DeviceScene> scene_;
  rad_render_desc desc_;
  rad_job* job_ = nullptr;
  std::vector<float> image_;
  rad_job_stats stats_{};
  std::chrono::steady_clock::time_point t0_;
  bool waited_ = false;
};

class RendererFactory {
 public:
  virtual ~RendererFactory() = default;
  virtual std::string UniqueId() const = 0;
  virtual std::unique_ptr<Renderer> Create(std::shared_ptr<DeviceScene> scene) const = 0;
};

class FactoryManager {
 public:
  void RegisterFactory(std::unique_ptr<RendererFactory> f);
  const RendererFactory& GetFactory(const std::string& name) const;
  // "path" and "vol_path" on the HIP device.
  static FactoryManager Default();

 private:
  std::map<std::string, std::unique_ptr<RendererFactory>> factories_;
};

}  // namespace offline
}  // namespace rad
