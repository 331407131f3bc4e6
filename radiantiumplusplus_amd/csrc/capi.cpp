// extern "C" entry points of include/rad_hip.h: device context, scene upload,
// BVH build, ray-batch intersection, asynchronous render jobs and the
// per-function probes.  No exception crosses this boundary.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>

#include "../../include/rad_hip.h"
#include "bvh_gpu.h"
#include "render_job.h"
#include "scene_host.h"
#include "wavefront.h"

using namespace rad;

// Static staging for rad_hip_intersect_closest_noalloc: the rays and hits of
// one call, kNoAllocMaxRays at most (8 MiB of rays + 5 MiB of hits).  Each
// device holds its own copy until the process ends; its device address comes
// from hipGetSymbolAddress, resolved once per device by rad_hip_init.
constexpr uint32_t kNoAllocMaxRays = 1u << 18;
constexpr size_t kScratchRayBytes = size_t(kNoAllocMaxRays) * sizeof(rad_ray);
constexpr size_t kScratchBytes = kScratchRayBytes + size_t(kNoAllocMaxRays) * sizeof(rad_hit);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
thread_local std::string g_last_error;

rad_status Fail(rad_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

template <class F>
rad_status Guard(F&& f) {
  try {
    f();
    return RAD_OK;
  } catch (const StatusError& e) {
    return Fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return Fail(RAD_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    std::string m = e.what();
    rad_status code = m.find("hip") != std::string::npos ? RAD_ERR_HIP : RAD_ERR_INVALID_ARG;
    return Fail(code, m);
  }
}
}  // namespace

namespace rad {
constexpr int kDefaultBuilder = RAD_BVH_DEVICE;
}  // namespace rad

namespace rad {

void SetLastError(const std::string& msg) { g_last_error = msg; }

void UploadScene(rad_scene* s) {
  hipStream_t st = s->ctx->stream;
  HostScene& h = s->host;
  s->d_shapes.upload(h.shapes.data(), h.shapes.size() * sizeof(ShapeRec), st);
  s->d_pos.upload(h.pos.data(), h.pos.size() * sizeof(float), st);
  s->d_nrm.upload(h.nrm.data(), h.nrm.size() * sizeof(float), st);
  s->d_uv.upload(h.uv.data(), h.uv.size() * sizeof(float), st);
  s->d_idx.upload(h.idx.data(), h.idx.size() * sizeof(uint32_t), st);
  {
    // the per-triangle shading records (DevScene::tris)
    std::vector<float> tris(h.idx.size() / 3 * 24, 0.f);
    for (const ShapeRec& sr : h.shapes) {
      if (sr.type != SHAPE_MESH) continue;
      for (uint32_t t = 0; t < sr.tri_count; t++) {
        const size_t k = size_t(sr.tri_offset) + t;
        float* o = tris.data() + 24 * k;
        for (int j = 0; j < 3; j++) {
          const size_t f = size_t(sr.vtx_offset) + h.idx[3 * k + j];
          for (int a = 0; a < 3; a++) {
            o[3 * j + a] = h.pos[3 * f + a];
            o[9 + 3 * j + a] = h.nrm[3 * f + a];
          }
          o[18 + 2 * j] = h.uv[2 * f];
          o[18 + 2 * j + 1] = h.uv[2 * f + 1];
        }
      }
    }
    s->d_tris.upload(tris.data(), tris.size() * sizeof(float), st);
    HipCheck(hipStreamSynchronize(st), "triangle records upload");  // the host vector is released next
  }
  s->d_bsdfs.upload(h.bsdfs.data(), h.bsdfs.size() * sizeof(BsdfRec), st);
  s->d_lights.upload(h.lights.data(), h.lights.size() * sizeof(LightRec), st);
  s->d_cdf.upload(h.light_cdf.data(), h.light_cdf.size() * sizeof(float), st);
  s->d_media.upload(h.media.data(), h.media.size() * sizeof(MediumRec), st);
  s->d_grid.upload(h.grid.data(), h.grid.size() * sizeof(float), st);
  s->d_tex.upload(h.textures.data(), h.textures.size() * sizeof(TexRec), st);
  s->d_texels.upload(h.texels.data(), h.texels.size() * sizeof(float), st);
  s->d_ldist.upload(h.light_dist.data(), h.light_dist.size() * sizeof(float), st);
  DevScene& d = s->dev;
  d = DevScene{};
  d.shapes = s->d_shapes.as<ShapeRec>();
  d.shape_count = uint32_t(h.shapes.size());
  d.pos = s->d_pos.as<float>();
  d.nrm = s->d_nrm.as<float>();
  d.uv = s->d_uv.as<float>();
  d.idx = s->d_idx.as<uint32_t>();
  d.tris = s->d_tris.as<float4>();
  d.bsdfs = s->d_bsdfs.as<BsdfRec>();
  d.lights = s->d_lights.as<LightRec>();
  d.light_count = uint32_t(h.lights.size());
  d.env_light = h.env_light;
  d.light_pdf = 1.0f / float(h.lights.size());
  d.tex_lights = 0;
  for (const LightRec& l : h.lights) d.tex_lights |= l.tex >= 0 ? 1u : 0u;
  d.tex_bsdfs = 0;
  for (const BsdfRec& b : h.bsdfs) d.tex_bsdfs |= b.tex_any ? 1u : 0u;
  d.light_cdf = s->d_cdf.as<float>();
  d.media = s->d_media.as<MediumRec>();
  d.global_medium = h.global_medium;
  d.grid_data = s->d_grid.as<float>();
  d.textures = s->d_tex.as<TexRec>();
  d.texels = s->d_texels.as<float>();
  d.light_dist = s->d_ldist.as<float>();
  d.cam = h.cam;
  // skybox world sphere (skybox.cpp:136-141, bounding_sphere.cpp:15-18)
  f3 lo = mk(h.world_lo[0], h.world_lo[1], h.world_lo[2]), hi = mk(h.world_hi[0], h.world_hi[1], h.world_hi[2]);
  f3 c = (lo + hi) * 0.5f;
  d.world_center = c;
  d.world_radius = std::max(kRayEps, norm(c - hi) * (1.f + kShadowEps));
  d.quirks = 1;
  HipCheck(hipStreamSynchronize(st), "scene upload");
}

}  // namespace rad

extern "C" {

const char* rad_hip_last_error(void) { return g_last_error.c_str(); }

size_t rad_abi_sizeof(const char* name) {
  if (!name) return 0;
  std::string n(name);
#define RAD_SZ(T) \
  if (n == #T) return sizeof(T);
  RAD_SZ(rad_shape_desc)
  RAD_SZ(rad_bsdf_desc)
  RAD_SZ(rad_light_desc)
  RAD_SZ(rad_medium_desc)
  RAD_SZ(rad_camera_desc)
  RAD_SZ(rad_scene_desc)
  RAD_SZ(rad_ray)
  RAD_SZ(rad_hit)
  RAD_SZ(rad_render_desc)
  RAD_SZ(rad_job_stats)
  RAD_SZ(rad_guide_record)
  RAD_SZ(rad_server_opts)
  RAD_SZ(rad_request_times)
  RAD_SZ(rad_server_stats)
  RAD_SZ(rad_bvh_report)
  RAD_SZ(rad_tune_report)
  RAD_SZ(rad_bvh_opts)
  RAD_SZ(rad_scene_info)
#undef RAD_SZ
  return 0;
}

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device ordinal: the address of `scratch`, null until resolved

// With `device` current: the device address of its copy of `scratch`.
void* ResolveScratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (size_t(device) >= g_scratch_ptr.size()) g_scratch_ptr.resize(size_t(device) + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    HipCheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw StatusError(RAD_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}
}  // namespace

rad_status rad_hip_init(int32_t device_ordinal, rad_ctx** out) {
  if (!out) return Fail(RAD_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  return Guard([&] {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) throw StatusError(RAD_ERR_HIP, "no HIP device available: " + HipErr(e));
    if (device_ordinal < 0 || device_ordinal >= count) throw StatusError(RAD_ERR_INVALID_ARG, "device ordinal out of range");
    HipCheck(hipSetDevice(device_ordinal), "hipSetDevice");
    auto* c = new rad_ctx();
    c->device = device_ordinal;
    hipDeviceProp_t prop;
    HipCheck(hipGetDeviceProperties(&prop, device_ordinal), "hipGetDeviceProperties");
    c->arch = prop.gcnArchName;
    c->cu_count = prop.multiProcessorCount;
    HipCheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    c->scratch = ResolveScratch(device_ordinal);
    *out = c;
  });
}

void rad_hip_shutdown(rad_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

rad_status rad_hip_scene_create(rad_ctx* ctx, const rad_scene_desc* desc, rad_scene** out) {
  if (!ctx || !desc || !out) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  *out = nullptr;
  return Guard([&] {
    HipCheck(hipSetDevice(ctx->device), "hipSetDevice");
    auto s = std::make_unique<rad_scene>();
    s->ctx = ctx;
    AssembleScene(*desc, s->host);
    UploadScene(s.get());
    *out = s.release();
  });
}

rad_status rad_hip_bvh_build(rad_scene* scene, const rad_bvh_opts* opts) {
  if (!scene) return Fail(RAD_ERR_INVALID_ARG, "scene is NULL");
  return Guard([&] {
    std::lock_guard<std::mutex> lk(scene->mu);
    HipCheck(hipSetDevice(scene->ctx->device), "hipSetDevice");
    int maxLeaf = opts && opts->max_leaf_size > 0 ? opts->max_leaf_size : 4;
    int builder = opts ? opts->builder : kDefaultBuilder;
    if (builder != RAD_BVH_DEVICE && builder != RAD_BVH_HOST) throw StatusError(RAD_ERR_INVALID_ARG, "unknown BVH builder");
    HostScene& h = scene->host;
    uint32_t n = uint32_t(h.prim_bounds.size() / 6);
    auto t0 = std::chrono::steady_clock::now();
    uint32_t nodeCount = 0, depth = 0, stackNeed = 0;
    float sah = 0.f;
    if (builder == RAD_BVH_DEVICE) {
      DeviceBvhStats ds;
      BuildBvh4Device(h.prim_bounds.data(), h.prim_records.data(), n, maxLeaf, scene->ctx->stream, scene->d_nodes,
                      scene->d_prims, ds);
      nodeCount = ds.node_count;
      depth = ds.depth;
      stackNeed = ds.stack_need;
      sah = ds.sah_cost;
    } else {
      BvhBuildResult r;
      BuildBvh4Host(h.prim_bounds.data(), n, maxLeaf, r);
      std::vector<float4> ordered(3 * size_t(n));
      for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) ordered[3 * size_t(i) + k] = h.prim_records[3 * size_t(r.order[i]) + k];
      scene->d_prims.upload(ordered.data(), ordered.size() * sizeof(float4), scene->ctx->stream);
      scene->d_nodes.upload(r.nodes.data(), r.nodes.size() * sizeof(BvhNode4), scene->ctx->stream);
      HipCheck(hipStreamSynchronize(scene->ctx->stream), "BVH upload");
      nodeCount = uint32_t(r.nodes.size());
      depth = r.depth;
      stackNeed = r.stack_need;
      sah = r.sah_cost;
    }
    // The traversal keeps kTraceStackEntries stack entries per ray; a deeper
    // need would drop subtrees (missed hits, light leaks), so such a tree is
    // refused.  The scene keeps no half-built tree.
    if (stackNeed > kTraceStackEntries) {
      scene->bvh_ready = false;
      throw StatusError(RAD_ERR_UNSUPPORTED, "BVH needs " + std::to_string(stackNeed) +
                                                 " traversal stack entries; the traversal holds " +
                                                 std::to_string(kTraceStackEntries));
    }
    QuantizeBvh4(scene->d_nodes, nodeCount, scene->ctx->stream, scene->d_qnodes);
    scene->builder = builder;
    scene->bvh_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    scene->dev.nodes = scene->d_nodes.as<BvhNode4>();
    scene->dev.qnodes = scene->d_qnodes.as<BvhNodeQ>();
    scene->dev.prims = scene->d_prims.as<float4>();
    scene->dev.prim_count = n;
    scene->dev.node_count = nodeCount;
    {
      // RAD_TRACE_POOL (1..64) shrinks the pooled primitive tests so that
      // tests can drive the path for entries beyond the pool; images do not
      // depend on it.
      const char* e = std::getenv("RAD_TRACE_POOL");
      const long v = e ? std::strtol(e, nullptr, 10) : long(kTracePool);
      scene->dev.pool_cap = uint32_t(std::min<long>(std::max<long>(v, 1), long(kTracePool)));
    }
    scene->bvh_depth = depth;
    scene->bvh_stack_need = stackNeed;
    scene->sah_cost = sah;
    scene->bvh_ready = true;
  });
}

static void RequireBvh(rad_scene* s) {
  if (!s->bvh_ready) throw StatusError(RAD_ERR_STATE, "rad_hip_bvh_build has not been called on this scene");
}

rad_status rad_hip_autotune(rad_scene* scene, const rad_render_desc* probe, rad_tune_report* rep) {
  if (!scene || !probe || !rep) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  return Guard([&] {
    RequireBvh(scene);
    std::memset(rep, 0, sizeof(*rep));
    const uint64_t samples = uint64_t(probe->width > 0 ? probe->width : scene->host.cam.width) *
                             uint64_t(probe->height > 0 ? probe->height : scene->host.cam.height) * probe->spp;
    auto trial = [&](uint32_t slots, uint32_t bpc, uint32_t tpc) {
      rad_render_desc d = *probe;
      d.path_slots = slots;
      d.blocks_per_cu = bpc;
      d.trace_blocks_per_cu = tpc;
      d.flags &= ~uint32_t(RAD_RENDER_PROFILE);
      rad_job* j = StartJob(scene, d, TakeSpareArena(scene));
      rad_job_stats st{};
      try {
        WaitJob(j, nullptr, nullptr, &st);
      } catch (...) {
        DestroyJob(j);
        throw;
      }
      DestroyJob(j);
      float ms = float(st.render_ms);
      if (rep->trials < 24) rep->trial[rep->trials++] = rad_tune_trial{slots, bpc, tpc, ms};
      return ms;
    };
    trial(4u << 20, 8, 8);  // warm-up (first-use costs)
    uint32_t bestSlots = 4u << 20, bestB = 8, bestT = 8;
    float best = 3.4e38f;
    for (uint32_t slots : {1u << 20, 2u << 20, 4u << 20, 8u << 20, 16u << 20, 32u << 20, 64u << 20}) {
      if (slots > 2 * samples && slots != (1u << 20)) continue;
      float ms = trial(slots, bestB, bestT);
      if (ms < best) best = ms, bestSlots = slots;
    }
    for (uint32_t b : {4u, 16u}) {
      float ms = trial(bestSlots, b, bestT);
      if (ms < best) best = ms, bestB = b;
    }
    for (uint32_t t : {2u, 4u, 16u}) {
      float ms = trial(bestSlots, bestB, t);
      if (ms < best) best = ms, bestT = t;
    }
    scene->tuned_slots = bestSlots;
    scene->tuned_blocks_per_cu = bestB;
    scene->tuned_trace_blocks_per_cu = bestT;
    TakeSpareArena(scene).reset();  // trial arenas reach ~16 GiB at 64 M slots: not kept
    rep->path_slots = bestSlots;
    rep->blocks_per_cu = bestB;
    rep->trace_blocks_per_cu = bestT;
    rep->best_ms = best;
  });
}

rad_status rad_hip_bvh_check(rad_scene* scene, rad_bvh_report* rep) {
  if (!scene || !rep) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  return Guard([&] {
    std::lock_guard<std::mutex> lk(scene->mu);
    if (!scene->bvh_ready) throw StatusError(RAD_ERR_STATE, "scene has no BVH");
    HipCheck(hipSetDevice(scene->ctx->device), "hipSetDevice");
    const uint32_t m = scene->dev.node_count, n = scene->dev.prim_count;
    std::vector<BvhNode4> nodes(m);
    std::vector<BvhNodeQ> qnodes(m);
    std::vector<float4> prims(3 * size_t(n));
    HipCheck(hipMemcpy(nodes.data(), scene->d_nodes.ptr, size_t(m) * sizeof(BvhNode4), hipMemcpyDeviceToHost), "nodes");
    HipCheck(hipMemcpy(qnodes.data(), scene->d_qnodes.ptr, size_t(m) * sizeof(BvhNodeQ), hipMemcpyDeviceToHost),
             "quantised nodes");
    HipCheck(hipMemcpy(prims.data(), scene->d_prims.ptr, prims.size() * sizeof(float4), hipMemcpyDeviceToHost), "prims");
    std::memset(rep, 0, sizeof(*rep));
    std::vector<uint8_t> seen(n, 0), visited(m, 0);
    std::vector<uint32_t> stack{0}, depthOf{1};
    auto inside = [](float a0, float a1, float b0, float b1) { return b0 <= a0 && a1 <= b1; };
    while (!stack.empty()) {
      uint32_t x = stack.back(), d = depthOf.back();
      stack.pop_back();
      depthOf.pop_back();
      if (x >= m || visited[x]) {
        rep->structure_errors++;
        continue;
      }
      visited[x] = 1;
      rep->depth = std::max(rep->depth, d);
      const BvhNode4& nd = nodes[x];
      {  // the traversal copy: same children and counts, dequantised boxes around the exact ones
        const BvhNodeQ& qn = qnodes[x];
        const float sx = quant_step(qn.exps), sy = quant_step(qn.exps >> 8), sz = quant_step(qn.exps >> 16);
        for (int s = 0; s < 4; s++) {
          if (qn.child[s] != nd.child[s]) {
            rep->quantized_box_violations++;
            continue;
          }
          if (nd.child[s] < 0) continue;
          const int sh = 8 * s;
          if (((qn.counts >> sh) & 0xffu) != nd.count[s] ||
              !(dequant((qn.qlox >> sh) & 0xffu, sx, qn.ox) <= nd.lox[s] && dequant((qn.qhix >> sh) & 0xffu, sx, qn.ox) >= nd.hix[s] &&
                dequant((qn.qloy >> sh) & 0xffu, sy, qn.oy) <= nd.loy[s] && dequant((qn.qhiy >> sh) & 0xffu, sy, qn.oy) >= nd.hiy[s] &&
                dequant((qn.qloz >> sh) & 0xffu, sz, qn.oz) <= nd.loz[s] && dequant((qn.qhiz >> sh) & 0xffu, sz, qn.oz) >= nd.hiz[s]))
            rep->quantized_box_violations++;
        }
      }
      for (int s = 0; s < 4; s++) {
        if (nd.child[s] < 0) continue;
        if (nd.count[s] > 0) {
          rep->max_leaf = std::max(rep->max_leaf, nd.count[s]);
          for (uint32_t k = 0; k < nd.count[s]; k++) {
            uint32_t p = uint32_t(nd.child[s]) + k;
            if (p >= n) {
              rep->structure_errors++;
              continue;
            }
            if (seen[p]++) rep->structure_errors++;
            const float4 a = prims[3 * size_t(p)], b = prims[3 * size_t(p) + 1], c = prims[3 * size_t(p) + 2];
            uint32_t kind;
            std::memcpy(&kind, &c.w, 4);
            float lo[3], hi[3];
            if (kind == PRIM_TRI) {
              lo[0] = std::min({a.x, b.x, c.x}), hi[0] = std::max({a.x, b.x, c.x});
              lo[1] = std::min({a.y, b.y, c.y}), hi[1] = std::max({a.y, b.y, c.y});
              lo[2] = std::min({a.z, b.z, c.z}), hi[2] = std::max({a.z, b.z, c.z});
            } else if (kind == PRIM_SPHERE) {
              lo[0] = a.x - a.w, hi[0] = a.x + a.w, lo[1] = a.y - a.w, hi[1] = a.y + a.w, lo[2] = a.z - a.w, hi[2] = a.z + a.w;
            } else {
              continue;  // rectangles: bounded through their shape transform
            }
            if (!(inside(lo[0], hi[0], nd.lox[s], nd.hix[s]) && inside(lo[1], hi[1], nd.loy[s], nd.hiy[s]) &&
                  inside(lo[2], hi[2], nd.loz[s], nd.hiz[s])))
              rep->leaf_box_violations++;
          }
        } else {
          uint32_t c = uint32_t(nd.child[s]);
          if (c >= m) {
            rep->structure_errors++;
            continue;
          }
          const BvhNode4& ch = nodes[c];
          for (int t = 0; t < 4; t++) {
            if (ch.child[t] < 0) continue;
            if (!(inside(ch.lox[t], ch.hix[t], nd.lox[s], nd.hix[s]) && inside(ch.loy[t], ch.hiy[t], nd.loy[s], nd.hiy[s]) &&
                  inside(ch.loz[t], ch.hiz[t], nd.loz[s], nd.hiz[s])))
              rep->child_box_violations++;
          }
          stack.push_back(c);
          depthOf.push_back(d + 1);
        }
      }
    }
    for (uint32_t p = 0; p < n; p++)
      if (seen[p] != 1) rep->unreferenced_or_repeated++;
    for (uint32_t x = 0; x < m; x++)
      if (!visited[x]) rep->structure_errors++;
    rep->node_count = m;
    rep->primitive_count = n;
    rep->sah_cost = BvhSahCost(nodes);
    rep->builder = scene->builder;
  });
}

rad_status rad_hip_scene_info(const rad_scene* scene, rad_scene_info* info) {
  if (!scene || !info) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  std::memset(info, 0, sizeof(*info));
  info->primitive_count = scene->host.prim_bounds.size() / 6;
  info->node_count = scene->dev.node_count;
  info->device_bytes = scene->device_bytes();
  info->bvh_build_ms = scene->bvh_ms;
  for (int a = 0; a < 3; a++) {
    info->world_min[a] = scene->host.world_lo[a];
    info->world_max[a] = scene->host.world_hi[a];
  }
  info->bvh_depth = scene->bvh_depth;
  info->sah_cost = scene->sah_cost;
  info->bvh_stack_need = scene->bvh_stack_need;
  return RAD_OK;
}

void rad_hip_scene_destroy(rad_scene* scene) {
  if (!scene) return;
  (void)hipSetDevice(scene->ctx->device);
  delete scene;
}


rad_status rad_hip_intersect_closest(rad_scene* scene, const rad_ray* rays, uint32_t n, rad_hit* hits) {
  if (!scene || (n && (!rays || !hits))) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  return Guard([&] {
    std::lock_guard<std::mutex> lk(scene->mu);
    RequireBvh(scene);
    HipCheck(hipSetDevice(scene->ctx->device), "hipSetDevice");
    hipStream_t st = scene->ctx->stream;
    scene->d_rays.alloc(std::max<size_t>(16, sizeof(rad_ray) * size_t(n)));
    scene->d_hits.alloc(std::max<size_t>(16, sizeof(rad_hit) * size_t(n)));
    if (n) HipCheck(hipMemcpyAsync(scene->d_rays.ptr, rays, sizeof(rad_ray) * size_t(n), hipMemcpyHostToDevice, st), "copy rays");
    LaunchClosest(scene->dev, scene->d_rays.ptr, n, scene->d_hits.ptr, st);
    HipCheck(hipGetLastError(), "closest-hit launch");
    if (n) HipCheck(hipMemcpyAsync(hits, scene->d_hits.ptr, sizeof(rad_hit) * size_t(n), hipMemcpyDeviceToHost, st), "copy hits");
    HipCheck(hipStreamSynchronize(st), "closest-hit");
  });
}

rad_status rad_hip_camera_footprint(rad_scene* scene, const float* in, uint32_t n, float* out) {
  if (!scene || (n && (!in || !out))) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  return Guard([&] {
    std::lock_guard<std::mutex> lk(scene->mu);
    RequireBvh(scene);
    HipCheck(hipSetDevice(scene->ctx->device), "hipSetDevice");
    hipStream_t st = scene->ctx->stream;
    DeviceBuffer din, dout;
    din.alloc(std::max<size_t>(16, size_t(n) * 8));
    dout.alloc(std::max<size_t>(16, size_t(n) * 24));
    if (n) HipCheck(hipMemcpyAsync(din.ptr, in, size_t(n) * 8, hipMemcpyHostToDevice, st), "copy positions");
    LaunchFootprint(scene->dev, din.as<float>(), n, dout.as<float>(), st);
    HipCheck(hipGetLastError(), "footprint launch");
    if (n) HipCheck(hipMemcpyAsync(out, dout.ptr, size_t(n) * 24, hipMemcpyDeviceToHost, st), "copy footprints");
    HipCheck(hipStreamSynchronize(st), "footprint");
  });
}

rad_status rad_hip_texture_eval(const rad_texture_desc* textures, uint32_t count, uint32_t index, const float* in,
                                uint32_t n, float* out) {
  if (!textures || index >= count || (n && (!in || !out))) return Fail(RAD_ERR_INVALID_ARG, "NULL argument or index out of range");
  return Guard([&] {
    int dev = 0;
    if (hipGetDeviceCount(&dev) != hipSuccess || dev == 0) throw StatusError(RAD_ERR_HIP, "no HIP device available");
    std::vector<TexRec> recs;
    std::vector<float> texels;
    BuildTextureSet(textures, count, recs, texels);
    DeviceBuffer dt, dx, din, dout;
    dt.upload(recs.data(), recs.size() * sizeof(TexRec), nullptr);
    dx.upload(texels.data(), texels.size() * sizeof(float), nullptr);
    din.alloc(std::max<size_t>(16, size_t(n) * 24));
    dout.alloc(std::max<size_t>(16, size_t(n) * 12));
    if (n) HipCheck(hipMemcpy(din.ptr, in, size_t(n) * 24, hipMemcpyHostToDevice), "copy lookups");
    DevScene d{};
    d.textures = dt.as<TexRec>();
    d.texels = dx.as<float>();
    LaunchTexEval(d, index, din.as<float>(), n, dout.as<float>(), nullptr);
    HipCheck(hipGetLastError(), "texture probe launch");
    if (n) HipCheck(hipMemcpy(out, dout.ptr, size_t(n) * 12, hipMemcpyDeviceToHost), "copy results");
  });
}

rad_status rad_hip_intersect_any(rad_scene* scene, const rad_ray* rays, uint32_t n, uint8_t* occluded) {
  if (!scene || (n && (!rays || !occluded))) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  return Guard([&] {
    std::lock_guard<std::mutex> lk(scene->mu);
    RequireBvh(scene);
    HipCheck(hipSetDevice(scene->ctx->device), "hipSetDevice");
    hipStream_t st = scene->ctx->stream;
    scene->d_rays.alloc(std::max<size_t>(16, sizeof(rad_ray) * size_t(n)));
    scene->d_occ.alloc(std::max<size_t>(16, size_t(n)));
    if (n) HipCheck(hipMemcpyAsync(scene->d_rays.ptr, rays, sizeof(rad_ray) * size_t(n), hipMemcpyHostToDevice, st), "copy rays");
    LaunchAny(scene->dev, scene->d_rays.ptr, n, scene->d_occ.as<uint8_t>(), st);
    HipCheck(hipGetLastError(), "occlusion launch");
    if (n) HipCheck(hipMemcpyAsync(occluded, scene->d_occ.ptr, size_t(n), hipMemcpyDeviceToHost, st), "copy results");
    HipCheck(hipStreamSynchronize(st), "occlusion");
  });
}

// rad_hip_intersect_closest staged through the static array instead of
// device buffers the library allocates: no hipMalloc / hipFree on any path.
rad_status rad_hip_intersect_closest_noalloc(rad_scene* scene, const rad_ray* rays, uint32_t n, rad_hit* hits) {
  if (!scene || (n && (!rays || !hits))) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  if (n > kNoAllocMaxRays)
    return Fail(RAD_ERR_INVALID_ARG, "rad_hip_intersect_closest_noalloc: " + std::to_string(n) +
                                         " rays exceed its static staging capacity of " +
                                         std::to_string(kNoAllocMaxRays) + " rays per call");
  return Guard([&] {
    std::lock_guard<std::mutex> lk(scene->mu);
    RequireBvh(scene);
    HipCheck(hipSetDevice(scene->ctx->device), "hipSetDevice");
    char* base = static_cast<char*>(scene->ctx->scratch);
    if (!base) throw StatusError(RAD_ERR_STATE, "static staging of this device was not resolved");
    rad_ray* d_rays = reinterpret_cast<rad_ray*>(base);
    rad_hit* d_hits = reinterpret_cast<rad_hit*>(base + kScratchRayBytes);
    hipStream_t st = scene->ctx->stream;
    if (n) HipCheck(hipMemcpyAsync(d_rays, rays, sizeof(rad_ray) * size_t(n), hipMemcpyHostToDevice, st), "copy rays");
    LaunchClosest(scene->dev, d_rays, n, d_hits, st);
    HipCheck(hipGetLastError(), "closest-hit launch");
    if (n) HipCheck(hipMemcpyAsync(hits, d_hits, sizeof(rad_hit) * size_t(n), hipMemcpyDeviceToHost, st), "copy hits");
    HipCheck(hipStreamSynchronize(st), "closest-hit");
  });
}

rad_status rad_hip_intersect_closest_device(rad_scene* scene, const rad_ray* d_rays, uint32_t n, rad_hit* d_hits,
                                            void* stream) {
  if (!scene || (n && (!d_rays || !d_hits))) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  return Guard([&] {
    RequireBvh(scene);
    LaunchClosest(scene->dev, d_rays, n, d_hits, (hipStream_t)stream);
    HipCheck(hipGetLastError(), "closest-hit launch");
  });
}

rad_status rad_hip_intersect_any_device(rad_scene* scene, const rad_ray* d_rays, uint32_t n, uint8_t* d_occ,
                                        void* stream) {
  if (!scene || (n && (!d_rays || !d_occ))) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  return Guard([&] {
    RequireBvh(scene);
    LaunchAny(scene->dev, d_rays, n, d_occ, (hipStream_t)stream);
    HipCheck(hipGetLastError(), "occlusion launch");
  });
}

rad_status rad_hip_render_start(rad_scene* scene, const rad_render_desc* desc, rad_job** out) {
  if (!scene || !desc || !out) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  *out = nullptr;
  return Guard([&] {
    RequireBvh(scene);
    rad_render_desc d = *desc;
    ResolveLaunch(scene, d);
    *out = StartJob(scene, d, TakeSpareArena(scene));
  });
}

rad_status rad_hip_job_poll(rad_job* job, uint64_t* done, uint64_t* total, int32_t* complete) {
  if (!job) return Fail(RAD_ERR_INVALID_ARG, "job is NULL");
  PollJob(job, done, total, complete);
  return RAD_OK;
}

rad_status rad_hip_job_cancel(rad_job* job) {
  if (!job) return Fail(RAD_ERR_INVALID_ARG, "job is NULL");
  CancelJob(job);
  return RAD_OK;
}

rad_status rad_hip_job_wait(rad_job* job, float* fb_rgb, float* fb_sumsq, rad_job_stats* stats) {
  if (!job) return Fail(RAD_ERR_INVALID_ARG, "job is NULL");
  return Guard([&] { WaitJob(job, fb_rgb, fb_sumsq, stats); });
}

void rad_hip_job_destroy(rad_job* job) { DestroyJob(job); }

static rad_status BsdfProbe(const rad_bsdf_desc* bsdf, uint32_t quirks, const float* in, uint32_t n, float* out,
                            bool sample) {
  if (!bsdf || (n && (!in || !out))) return Fail(RAD_ERR_INVALID_ARG, "NULL argument");
  return Guard([&] {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw StatusError(RAD_ERR_HIP, "no HIP device available");
    if (bsdf->type == RAD_BSDF_TWO_SIDE || bsdf->type == RAD_BSDF_MASK)
      throw StatusError(RAD_ERR_UNSUPPORTED, "the probes take a base BSDF (two_side / mask wrap another record)");
    BsdfRec b = MakeBsdfRec(*bsdf);
    size_t inBytes = size_t(n) * 6 * sizeof(float), outBytes = size_t(n) * (sample ? 8 : 4) * sizeof(float);
    DeviceBuffer din, dout;
    din.alloc(std::max<size_t>(inBytes, 16));
    dout.alloc(std::max<size_t>(outBytes, 16));
    HipCheck(hipMemcpy(din.ptr, in, inBytes, hipMemcpyHostToDevice), "copy probe input");
    if (sample)
      LaunchBsdfSample(b, quirks != 0, din.as<float>(), n, dout.as<float>(), nullptr);
    else
      LaunchBsdfEval(b, quirks != 0, din.as<float>(), n, dout.as<float>(), nullptr);
    HipCheck(hipGetLastError(), "probe launch");
    HipCheck(hipMemcpy(out, dout.ptr, outBytes, hipMemcpyDeviceToHost), "copy probe output");
  });
}

rad_status rad_hip_bsdf_eval(const rad_bsdf_desc* bsdf, uint32_t quirks, const float* in, uint32_t n, float* out) {
  return BsdfProbe(bsdf, quirks, in, n, out, false);
}

rad_status rad_hip_bsdf_sample(const rad_bsdf_desc* bsdf, uint32_t quirks, const float* in, uint32_t n, float* out) {
  return BsdfProbe(bsdf, quirks, in, n, out, true);
}

}  // extern "C"
