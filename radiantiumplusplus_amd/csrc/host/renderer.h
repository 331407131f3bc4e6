// C++ host interface over the C ABI, shaped like the reference's renderer
// plugin API so a reference-side caller can switch to the GPU path:
//   Factory / FactoryManager ~ core/factory.h:13-61, core/factory.cpp:5-19: a
//                    factory carries its name from construction; duplicate
//                    and unknown names throw
//   RendererFactory ~ build/factory.h:84-89: Create(BuildContext*, scene, config)
//   BuildContext    ~ build/build_context.h:20-58: SetFromJson, the default
//                    factory manager, Build() -> the renderer named by
//                    renderer.type (build_context.cpp:88-321)
//   Renderer        ~ render/renderer.h:19-48: owns its scene; Start (async),
//                    Wait, Stop, IsComplete, AllTaskCount, CompletedTaskCount,
//                    ElapsedTime (ms), SaveResult (<name>.exr)
// The factories "path" and "vol_path" build GPU renderers; renderer.device
// "hip" is the default and the only device this library provides.
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

// A scene loaded from JSON and resident on the device (the Scene that
// BuildContext::Build assembles, with the BVH in place of the Embree accel).
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

class BuildContext;

class Renderer {
 public:
  Renderer(BuildContext* ctx, std::unique_ptr<DeviceScene> scene, const rad_render_desc& cfg);
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
  std::unique_ptr<DeviceScene> scene_;
  rad_render_desc desc_;
  rad_job* job_ = nullptr;
  std::vector<float> image_;
  rad_job_stats stats_{};
  std::chrono::steady_clock::time_point t0_;
  bool waited_ = false;
};

class Factory {
 public:
  explicit Factory(std::string name) : name_(std::move(name)) {}
  virtual ~Factory() = default;
  const std::string& GetName() const { return name_; }

 protected:
  std::string name_;
};

class RendererFactory : public Factory {
 public:
  explicit RendererFactory(std::string name) : Factory(std::move(name)) {}
  // cfg: the renderer settings of the scene (renderer.* and sampler.*)
  virtual std::unique_ptr<Renderer> Create(BuildContext* ctx, std::unique_ptr<DeviceScene> scene,
                                           const rad_render_desc& cfg) const = 0;
};

class FactoryManager {
 public:
  void RegisterFactory(std::unique_ptr<Factory> f);
  Factory* GetFactory(const std::string& name) const;
  template <typename T>
  T* GetFactory(const std::string& name) const {
    return dynamic_cast<T*>(GetFactory(name));
  }

 private:
  std::map<std::string, std::unique_ptr<Factory>> factories_;
};

// "path" and "vol_path" on the HIP device (GetRadOfflineFactories, build/factory.cpp:39-75).
void RegisterHipRenderers(FactoryManager& m);

class BuildContext {
 public:
  explicit BuildContext(rad_ctx* ctx) : ctx_(ctx) {}
  void SetFactoryManager(FactoryManager& mgr) { mgr_ = &mgr; }
  void SetDefaultFactoryManager();
  const FactoryManager& GetFactoryManager() const;
  // The scene file (build_context.cpp:47-76 reads the same keys).
  void SetFromJson(const std::string& path) { path_ = path; }
  // Optional overrides of sampler.sample_count and of the guiding mode.
  void SetSampleCount(uint32_t spp) { spp_ = spp; }
  void SetGuiding(int32_t mode) { guiding_ = mode; }
  // Loads the scene onto the device and creates the renderer named by
  // renderer.type (build_context.cpp:317-320).
  std::unique_ptr<Renderer> Build();
  // BVH build time of the last Build (embree.cpp:38-42 logs the same).
  double LastBvhBuildMs() const { return bvh_ms_; }

 private:
  rad_ctx* ctx_;
  FactoryManager* mgr_ = nullptr;
  std::unique_ptr<FactoryManager> default_;
  std::string path_;
  uint32_t spp_ = 0;
  int32_t guiding_ = -1;
  double bvh_ms_ = 0;
};

}  // namespace offline
}  // namespace rad
