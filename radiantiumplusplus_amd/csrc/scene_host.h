// Host-side scene assembly: turns a rad_scene_desc into the device records of
// scene_dev.h, builds the BVH and uploads everything to HBM.  Mirrors the
// object construction of BuildContext::Build (build_context.cpp:88-321) and
// the shape / light / medium constructors it calls.
#pragma once

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rad_hip.h"
#include "bvh_build.h"
#include "scene_dev.h"

struct rad_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // device address of this device's static staging array (capi.cpp)
  std::string arch;
  int cu_count = 0;
};

namespace rad {

struct HostScene {
  std::vector<ShapeRec> shapes;
  std::vector<float> pos, nrm, uv;
  std::vector<uint32_t> idx;
  std::vector<BsdfRec> bsdfs;
  std::vector<LightRec> lights;
  std::vector<float> light_cdf;
  std::vector<float> light_dist;  // 2-D luminance distributions of textured emitters
  std::vector<MediumRec> media;
  std::vector<float> grid;
  std::vector<TexRec> textures;
  std::vector<float> texels;  // every bitmap's mip chain
  int32_t env_light = -1;
  int32_t global_medium = -1;
  CameraRec cam{};
  rad_camera_desc camera_desc{};
  // primitive list before BVH ordering
  std::vector<float> prim_bounds;       // 6 per primitive
  std::vector<float4> prim_records;     // 3 per primitive
  float world_lo[3], world_hi[3];
  bool quirks_default = true;
};

// Throws std::runtime_error on invalid input.
void AssembleScene(const rad_scene_desc& desc, HostScene& out);

CameraRec MakeCameraRec(const rad_camera_desc& c);

// Texture records and the texel pool (each bitmap's mip chain) of a texture array.
void BuildTextureSet(const rad_texture_desc* descs, uint32_t n, std::vector<TexRec>& textures, std::vector<float>& texels);

// Reference-default BSDF record from the desc (eta derivations of disney.cpp:86-101 etc.).
BsdfRec MakeBsdfRec(const rad_bsdf_desc& d);

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { reset(); }
  void reset();
  void alloc(size_t n);
  void upload(const void* src, size_t n, hipStream_t s);
  template <class T>
  T* as() const { return reinterpret_cast<T*>(ptr); }
};

std::string HipErr(hipError_t e);
void HipCheck(hipError_t e, const char* what);

}  // namespace rad

namespace rad {
namespace guide {
struct GuideState;
}
}  // namespace rad

struct rad_scene {
  rad_ctx* ctx = nullptr;
  rad::HostScene host;
  rad::DevScene dev{};
  rad::DeviceBuffer d_nodes, d_qnodes, d_prims, d_shapes, d_pos, d_nrm, d_uv, d_idx, d_tris, d_bsdfs, d_lights, d_cdf, d_media, d_grid,
      d_tex, d_texels, d_ldist;
  rad::DeviceBuffer d_rays, d_hits, d_occ;  // staging for the host-pointer intersect entry points
  std::mutex mu;
  bool bvh_ready = false;
  double bvh_ms = 0;
  uint32_t bvh_depth = 0;
  uint32_t bvh_stack_need = 0;
  int32_t builder = -1;
  float sah_cost = 0;
  std::shared_ptr<rad::guide::GuideState> guide;  // created on first guided render
  std::mutex guide_mu;  // held by a render that trains the guide
  // launch configuration chosen by rad_hip_autotune (0 = library default);
  // renders resolve it once, when they are submitted (ResolveLaunch)
  std::atomic<uint32_t> tuned_slots{0}, tuned_blocks_per_cu{0}, tuned_trace_blocks_per_cu{0};
  // The arena of the last finished render job, kept for the next one: fresh
  // device allocations are cleared by the driver in the background, which
  // competed with the first renders after each allocation.
  std::unique_ptr<rad::DeviceBuffer> spare_arena;
  std::mutex arena_mu;
  uint64_t device_bytes() const;
};
