// Device-resident scene layout.
//
// Geometry for traversal lives in two HBM arrays:
//   nodes  - 4-wide BVH nodes, 128 B each (one cache line): six 4-float slab
//            planes, four child links, four leaf counts;
//   prims  - primitives in leaf order, three float4 (48 B) each:
//            triangle: (v0, geomID) (v1, primID) (v2, PRIM_TRI)
//            sphere:   (center, radius) (-, geomID) (-, PRIM_SPHERE)
//            rectangle:(-) (-, geomID) (-, PRIM_RECT); matrix in its ShapeRec.
// Shading data (vertex attributes, per-shape records, BSDF / light / medium
// records) is separate, so traversal touches only nodes + prims.
#pragma once

#include "../../include/rad_hip.h"
#include "rad_math.h"

namespace rad {

enum : uint32_t { PRIM_TRI = 0, PRIM_SPHERE = 1, PRIM_RECT = 2 };
enum : uint32_t { SHAPE_HAS_NORMALS = 1, SHAPE_HAS_UVS = 2 };
enum : int32_t { SHAPE_MESH = 0, SHAPE_SPHERE = 1, SHAPE_RECTANGLE = 2 };

struct alignas(128) BvhNode4 {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];   // inner: node index; leaf: first primitive; empty: -1
  uint32_t count[4];  // 0: inner node; > 0: leaf primitive count
};
static_assert(sizeof(BvhNode4) == 128, "BVH node must be one 128-byte line");

// Entries of a traversal stack (node, entry distance).  A node visit pushes
// all but the nearest of its hit inner children, so a tree needs at most
// need(n) = (inner children of n - 1) + max over inner children c of need(c)
// entries (0 for a node whose children are all leaves); the builders compute
// need(root) and a tree that needs more than this is refused.
constexpr uint32_t kTraceStackEntries = 64;
// Entries of a wave's pooled primitive tests per node step (traverse.h); a
// lane whose entries fall beyond the pool tests them itself.
constexpr uint32_t kTracePool = 64;

// The traversal's copy of a BvhNode4 in 64 bytes: the node box origin, a
// power-of-two step per axis, and every child's bounds as 8-bit multiples of
// the step, rounded outward, so each child box contains the exact one.  A
// dequantised bound is fmaf(q, step, origin): one rounding, computed the same
// way when the node is built, where the outward rounding is verified.
struct alignas(64) BvhNodeQ {
  float ox, oy, oz;
  uint32_t exps;                    // biased exponents of the x, y, z steps (bytes 0, 1, 2)
  uint32_t qlox, qhix, qloy, qhiy;  // child i's bound in byte i
  uint32_t qloz, qhiz;
  uint32_t counts;                  // child i's leaf primitive count in byte i (0: inner)
  uint32_t pad;
  int32_t child[4];                 // as in BvhNode4
};
static_assert(sizeof(BvhNodeQ) == 64, "quantised BVH node must be 64 bytes");

__host__ __device__ inline float quant_step(uint32_t biased) {
  union {
    uint32_t u;
    float f;
  } x;
  x.u = (biased & 0xffu) << 23;
  return x.f;
}
__host__ __device__ inline float dequant(uint32_t q, float step, float origin) { return fmaf(float(q), step, origin); }

struct ShapeRec {
  int32_t type;       // SHAPE_*
  int32_t bsdf;
  int32_t light;
  int32_t med_in, med_out;
  uint32_t flags;     // SHAPE_HAS_*
  uint32_t vtx_offset;
  uint32_t tri_offset;
  uint32_t tri_count;
  float area;
  Affine to_world;    // rectangle: ToWorld; sphere: _toWorld (UV parameterisation)
  Affine to_local;    // rectangle: ToObject; sphere: inverse of _toWorld
  f3 frame_s, frame_t, frame_n;  // rectangle frame (dpdu, dpdv, N)
  f3 center;
  float radius;
  uint32_t cdf_offset;  // mesh emitters: triangle-area CDF in DevScene::light_cdf
  float cdf_sum, cdf_norm;
};

enum : int32_t {
  BSDF_DIFFUSE = 0,
  BSDF_ROUGH_METAL = 1,
  BSDF_ROUGH_GLASS = 2,
  BSDF_DISNEY = 3,
  BSDF_PLASTIC = 4,
  BSDF_ROUGH_PLASTIC = 5,
  BSDF_PERFECT_MIRROR = 6,
  BSDF_PERFECT_GLASS = 7,
  BSDF_TWO_SIDE = 8,  // wraps bsdfs[nested]
  BSDF_MASK = 9       // wraps bsdfs[nested]
};
enum : uint32_t {
  LOBE_DIFFUSE = 1,
  LOBE_GLOSSY = 2,
  LOBE_DELTA = 4,
  LOBE_REFLECTION = 8,
  LOBE_TRANSMISSION = 16,
  LOBE_NO_DELTA = LOBE_DIFFUSE | LOBE_GLOSSY
};

struct BsdfRec {
  int32_t type;
  uint32_t lobes;         // BsdfType flags of the instance (render/bsdf.h:10-40)
  int32_t beckmann;
  int32_t sample_visible;
  uint32_t disney_flags;
  f3 reflectance, transmittance, eta3, k3;
  float alpha_u, alpha_v;
  float eta, inv_eta;     // rough_glass int/ext; disney eta
  f3 base_color;
  float roughness, anisotropic, spec_trans, sheen, sheen_tint, flatness, spec_tint, metallic, clearcoat,
      clearcoat_gloss;
  float spec_rate, cc_rate, diff_rate;
  // plastic / rough_plastic (plastic.cpp:20-40): colours, 1 / eta^2, the
  // internal Fresnel average Fdr(1 / eta), the specular sampling weight
  f3 diffuse3, spec3;
  float inv_eta2, fdr_int, spec_weight;
  int32_t nonlinear;
  // two_side / mask
  int32_t nested;
  float opacity;
  // textured parameters: bit k of tex_mask set = slot k (RAD_TEX_*) reads
  // textures[tex[k]]; tex_any: this record or a record it wraps is textured
  uint32_t tex_mask;
  uint32_t tex_any;
  uint16_t tex[20];
};

// A texture (texture/bitmap.cpp, texture/chessboard.cpp).  Bitmap texels are
// float, `channels` per texel, level l at texels + off[l], row-major.
constexpr int kMaxMips = 16;
enum : int32_t { TEX_BITMAP = 0, TEX_CHESSBOARD = 1 };
struct TexRec {
  int32_t type, channels, filter, wrap;
  uint32_t levels;
  float aniso;
  uint32_t off[kMaxMips];
  uint32_t w[kMaxMips], h[kMaxMips];
  // chessboard: to_uv (affine with the homogeneous divide), nested bitmaps
  // (-1 = the constant colour), remap
  float to_uv[16];
  int32_t tex_a, tex_b;
  f3 color_a, color_b;
  int32_t remap;
};

enum : int32_t { LIGHT_AREA = 0, LIGHT_SKYBOX = 1, LIGHT_POINT = 2, LIGHT_PROJECTION = 3 };
struct LightRec {
  int32_t type;
  int32_t shape;
  f3 radiance;      // area / skybox radiance, point intensity, projection irradiance x scale
  Affine to_world;  // skybox orientation; point position (translation); projection frame
  Affine to_local;  // projection: world -> light frame
  float clip[16];   // projection: light frame -> [0,1]^2 image plane (row-major, projective)
  f3 axis;          // projection: normalised +z of the frame in world space
  // radiance texture (area / skybox; -1 = constant).  A texture of more than
  // one texel is sampled through a 2-D luminance distribution at
  // light_dist + dist_off: lum[h][w] | cdf[h][w + 1] | 1/row sum[h] |
  // row sum[h] | marginal cdf[h + 1]; msum = the marginal's sum.
  int32_t tex;
  int32_t tex_uniform;
  uint32_t dist_off, dist_w, dist_h;
  float msum, mnorm;
};

enum : int32_t { MEDIUM_HOMOGENEOUS = 0, MEDIUM_HETEROGENEOUS = 1 };
struct MediumRec {
  int32_t type;
  int32_t phase;      // 0 isotropic, 1 HG
  float g;
  f3 sigma_t;         // homogeneous sigma_t * scale (or constant heterogeneous)
  f3 albedo;
  float scale;
  float max_density;  // heterogeneous majorant (grid max * scale)
  Affine to_local;    // world -> [0,1]^3 grid space
  f3 bbox_lo, bbox_hi;
  int32_t sx, sy, sz, channels, wrap;
  int32_t has_grid;
  uint64_t grid_offset;  // float offset into DevScene::grid_data
};

struct CameraRec {
  Mat4 clip_to_camera;   // inverse of scale * translate * perspective (perspective.cpp:33-52)
  Affine camera_to_world;
  float near_clip, far_clip;
  int32_t width, height;
  float rcp_w, rcp_h;
  f3 dx, dy;             // camera-space offset of one pixel on the near plane (perspective.cpp:55-58)
};

struct DevScene {
  const BvhNode4* nodes;   // full-precision nodes (checks, statistics)
  const BvhNodeQ* qnodes;  // traversal copy
  const float4* prims;
  uint32_t prim_count, node_count;
  const ShapeRec* shapes;
  uint32_t shape_count;
  const float* pos;        // world-space positions, 3 floats per vertex
  const float* nrm;        // world-space (unnormalised) normals, 3 floats per vertex
  const float* uv;         // 2 floats per vertex
  const uint32_t* idx;     // 3 indices per triangle, local to the shape's vertex range
  // Per triangle (tri_offset + prim), 6 float4: its positions p0 p1 p2, normals
  // n0 n1 n2 and uvs, packed in that order (24 floats; zeros where the mesh has
  // none), so shading a hit reads one 96-byte record instead of the indices
  // and nine scattered vertex attributes.
  const float4* tris;
  const BsdfRec* bsdfs;
  const LightRec* lights;
  uint32_t light_count;
  int32_t env_light;
  float light_pdf;
  uint32_t tex_lights;  // 1: some light's radiance (irradiance) is a texture
  uint32_t tex_bsdfs;   // 1: some BSDF has a textured parameter
  const float* light_cdf;
  const MediumRec* media;
  int32_t global_medium;
  const float* grid_data;
  const TexRec* textures;
  const float* texels;
  const float* light_dist;
  CameraRec cam;
  f3 world_center;
  float world_radius;
  uint32_t quirks;
  uint32_t pool_cap;  // pooled primitive tests per wave step, 1..kTracePool (smaller only to test the overflow path)
};

}  // namespace rad
