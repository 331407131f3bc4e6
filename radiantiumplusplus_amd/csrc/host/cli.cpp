// Command-line renderer, the counterpart of rad.offline.cli (main.cpp:8-71):
//   rad_offline_cli --scene <scene.json> [--out <dir>] [--spp N] [--guiding 0|1|2] [--device N]
// Loads the scene, builds the BVH, runs the renderer named by renderer.type
// ("path" / "vol_path") with a one-second progress line, and writes
// <scene name>.exr next to the scene (or into --out).
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <string>
#include <thread>

#include "renderer.h"

using namespace rad::offline;

int main(int argc, char** argv) {
  std::string scenePath, outDir;
  int spp = 0, guiding = -1, device = 0;
  for (int i = 1; i < argc; i++) {
    std::string a(argv[i]);
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
      return argv[++i];
    };
    try {
      if (a == "--scene")
        scenePath = next();
      else if (a == "--out")
        outDir = next();
      else if (a == "--spp")
        spp = std::stoi(next());
      else if (a == "--guiding")
        guiding = std::stoi(next());
      else if (a == "--device")
        device = std::stoi(next());
    } catch (const std::exception& e) {
      std::fprintf(stderr, "[error] %s\n", e.what());
      return 2;
    }
  }
  if (scenePath.empty()) {
    std::fprintf(stderr, "[error] should input cmd like \"--scene <scene.json>\"\n");
    return 2;
  }
  namespace fs = std::filesystem;
  fs::path p(scenePath);
  if (!fs::exists(p)) {
    std::fprintf(stderr, "[error] cannot open file: %s\n", scenePath.c_str());
    return 2;
  }
  rad_ctx* ctx = nullptr;
  int rc = 0;
  try {
    Check(rad_hip_init(device, &ctx), "init");
    BuildContext build(ctx);
    build.SetFromJson(p.string());
    build.SetDefaultFactoryManager();
    if (spp > 0) build.SetSampleCount(uint32_t(spp));
    if (guiding >= 0) build.SetGuiding(guiding);
    std::unique_ptr<Renderer> renderer = build.Build();
    std::printf("[info] build done. %.1f ms\n", build.LastBvhBuildMs());
    std::printf("[info] start rendering...\n");
    std::fflush(stdout);
    renderer->Start();
    std::atomic<bool> done{false};
    std::thread bar([&] {
      while (!done.load() && !renderer->IsComplete()) {
        uint64_t all = renderer->AllTaskCount(), cur = renderer->CompletedTaskCount();
        std::printf("\r[%5.1f%%] %llu / %llu samples, %lld ms", all ? 100.0 * double(cur) / double(all) : 0.0,
                    (unsigned long long)cur, (unsigned long long)all, (long long)renderer->ElapsedTime());
        std::fflush(stdout);
        std::this_thread::sleep_for(std::chrono::seconds(1));
      }
    });
    renderer->Wait();
    done.store(true);
    bar.join();
    std::printf("\n");
    std::string dir = outDir.empty() ? p.parent_path().string() : outDir;
    renderer->SaveResult(dir, p.filename().replace_extension().string());
    std::printf("[info] done. render used time: %lld ms\n", (long long)renderer->ElapsedTime());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "[error] unhandled exception: %s\n", e.what());
    rc = 1;
  }
  if (ctx) rad_hip_shutdown(ctx);
  return rc;
}
