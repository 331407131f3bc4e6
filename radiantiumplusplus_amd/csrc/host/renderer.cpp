#include "renderer.h"

#include <thread>

namespace rad {
namespace offline {

void Check(rad_status s, const char* what) {
  if (s != RAD_OK) throw RadException(s, std::string(what) + ": " + rad_hip_last_error());
}

DeviceScene::DeviceScene(rad_ctx* ctx, const std::string& jsonPath) {
  Check(rad_scene_load_json(jsonPath.c_str(), &loaded_), "load scene");
  render_ = rad_loaded_scene_render(loaded_);
  if (device() != "hip") {
    std::string d = device();
    rad_loaded_scene_free(loaded_);
    throw RadException(RAD_ERR_UNSUPPORTED, "renderer.device \"" + d + "\" is not provided by this library");
  }
  rad_status s = rad_hip_scene_create(ctx, rad_loaded_scene_desc(loaded_), &scene_);
  if (s == RAD_OK) s = rad_hip_bvh_build(scene_, nullptr);
  if (s != RAD_OK) {
    std::string m = rad_hip_last_error();
    if (scene_) rad_hip_scene_destroy(scene_);
    rad_loaded_scene_free(loaded_);
    throw RadException(s, "create scene: " + m);
  }
}

DeviceScene::~DeviceScene() {
  if (scene_) rad_hip_scene_destroy(scene_);
  if (loaded_) rad_loaded_scene_free(loaded_);
}

std::string DeviceScene::renderer_type() const { return rad_loaded_scene_renderer_type(loaded_); }
std::string DeviceScene::device() const { return rad_loaded_scene_device(loaded_); }
int DeviceScene::width() const { return rad_loaded_scene_desc(loaded_)->camera.width; }
int DeviceScene::height() const { return rad_loaded_scene_desc(loaded_)->camera.height; }
double DeviceScene::bvh_build_ms() const {
  rad_scene_info info{};
  Check(rad_hip_scene_info(scene_, &info), "scene info");
  return info.bvh_build_ms;
}

Renderer::Renderer(BuildContext*, std::unique_ptr<DeviceScene> scene, const rad_render_desc& cfg)
    : scene_(std::move(scene)), desc_(cfg) {}

Renderer::~Renderer() {
  if (job_) rad_hip_job_destroy(job_);
}

void Renderer::Start() {
  if (job_) throw RadException(RAD_ERR_STATE, "renderer already started");
  t0_ = std::chrono::steady_clock::now();
  Check(rad_hip_render_start(scene_->handle(), &desc_, &job_), "start render");
}

void Renderer::Wait() {
  if (!job_) throw RadException(RAD_ERR_STATE, "renderer not started");
  if (waited_) return;
  image_.assign(size_t(scene_->width()) * scene_->height() * 3, 0.f);
  rad_status s = rad_hip_job_wait(job_, image_.data(), nullptr, &stats_);
  waited_ = true;
  if (s != RAD_OK && s != RAD_ERR_CANCELLED) Check(s, "render");
}

void Renderer::Stop() {
  if (job_) Check(rad_hip_job_cancel(job_), "stop render");
}

bool Renderer::IsComplete() const {
  if (!job_) return false;
  uint64_t done = 0, total = 0;
  int32_t complete = 0;
  Check(rad_hip_job_poll(job_, &done, &total, &complete), "poll");
  return complete != 0;
}

uint64_t Renderer::AllTaskCount() const {
  uint64_t done = 0, total = 0;
  int32_t complete = 0;
  if (job_) Check(rad_hip_job_poll(job_, &done, &total, &complete), "poll");
  return total;
}

uint64_t Renderer::CompletedTaskCount() const {
  uint64_t done = 0, total = 0;
  int32_t complete = 0;
  if (job_) Check(rad_hip_job_poll(job_, &done, &total, &complete), "poll");
  return done;
}

int64_t Renderer::ElapsedTime() const {
  return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0_).count();
}

void Renderer::SaveResult(const std::string& directory, const std::string& name) const {
  if (image_.empty()) throw RadException(RAD_ERR_STATE, "no result (call Wait first)");
  std::string path = (directory.empty() ? std::string(".") : directory) + "/" + name + ".exr";
  Check(rad_write_exr(path.c_str(), image_.data(), scene_->width(), scene_->height()), "save result");
}

namespace {

class HipRendererFactory : public RendererFactory {
 public:
  HipRendererFactory(const char* name, int32_t integrator) : RendererFactory(name), integrator_(integrator) {}
  std::unique_ptr<Renderer> Create(BuildContext* ctx, std::unique_ptr<DeviceScene> scene,
                                   const rad_render_desc& cfg) const override {
    rad_render_desc d = cfg;
    d.integrator = integrator_;
    return std::make_unique<Renderer>(ctx, std::move(scene), d);
  }

 private:
  int32_t integrator_;
};

}  // namespace

void FactoryManager::RegisterFactory(std::unique_ptr<Factory> f) {
  const std::string name = f->GetName();
  if (factories_.count(name)) throw RadException(RAD_ERR_INVALID_ARG, "factory already registered: " + name);
  factories_[name] = std::move(f);
}

Factory* FactoryManager::GetFactory(const std::string& name) const {
  auto it = factories_.find(name);
  if (it == factories_.end()) throw RadException(RAD_ERR_UNSUPPORTED, "no factory named " + name);
  return it->second.get();
}

void RegisterHipRenderers(FactoryManager& m) {
  m.RegisterFactory(std::make_unique<HipRendererFactory>("path", RAD_INTEGRATOR_PATH));
  m.RegisterFactory(std::make_unique<HipRendererFactory>("vol_path", RAD_INTEGRATOR_VOL_PATH));
}

void BuildContext::SetDefaultFactoryManager() {
  default_ = std::make_unique<FactoryManager>();
  RegisterHipRenderers(*default_);
  mgr_ = default_.get();
}

const FactoryManager& BuildContext::GetFactoryManager() const {
  if (!mgr_) throw RadException(RAD_ERR_STATE, "no factory manager set");
  return *mgr_;
}

std::unique_ptr<Renderer> BuildContext::Build() {
  if (path_.empty()) throw RadException(RAD_ERR_STATE, "no scene set (SetFromJson)");
  auto scene = std::make_unique<DeviceScene>(ctx_, path_);
  bvh_ms_ = scene->bvh_build_ms();
  rad_render_desc cfg = scene->render_desc();
  if (spp_ > 0) cfg.spp = spp_;
  if (guiding_ >= 0) cfg.guiding = guiding_;
  const std::string type = scene->renderer_type();
  auto* factory = GetFactoryManager().GetFactory<RendererFactory>(type);
  if (!factory) throw RadException(RAD_ERR_UNSUPPORTED, "factory " + type + " is not a renderer factory");
  return factory->Create(this, std::move(scene), cfg);
}

}  // namespace offline
}  // namespace rad
