/*
 * rad_hip.h - C ABI of the MI355X-native `path` / `vol_path` integrators and
 * the scene intersection they call.
 *
 * This is the drop-in boundary for the renderer plugin of RadiantiumPlusPlus:
 *   - RendererFactory("path") / RendererFactory("vol_path")
 *       module/rad.offline/include/rad/offline/build/factory.h:84-89,
 *       module/rad.offline/src/renderer/path.cpp:243-254,
 *       module/rad.offline/src/renderer/vol_path.cpp:356-367
 *   - the Renderer contract Start/Wait/Stop/IsComplete/AllTaskCount/
 *     CompletedTaskCount/ElapsedTime/SaveResult
 *       module/rad.offline/include/rad/offline/render/renderer.h:19-48
 *   - the Accel plugin (closest hit / occlusion, world bound)
 *       module/rad.offline/include/rad/offline/render/accel.h:10-30,
 *       module/rad.offline/src/accel/embree.cpp:70-111
 *   - the scene build that feeds them (shapes, BSDFs, lights, media, camera)
 *       module/rad.offline/src/build/build_context.cpp:88-321
 *
 * Conventions: plain C types only, no exceptions cross the ABI, every call that
 * can fail returns rad_status (0 = OK) and leaves a thread-local message in
 * rad_hip_last_error().  Host buffers are owned by the caller; a scene handle
 * owns its device copies.  Matrices are row-major 4x4 (the reference's
 * `to_world` JSON layout, core/config_node.h:36-55).
 */
#ifndef RAD_HIP_H_
#define RAD_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rad_status;

enum {
  RAD_OK = 0,
  RAD_ERR_INVALID_ARG = 1,
  RAD_ERR_HIP = 2,
  RAD_ERR_OUT_OF_MEMORY = 3,
  RAD_ERR_UNSUPPORTED = 4,
  RAD_ERR_IO = 5,
  RAD_ERR_CANCELLED = 6,
  RAD_ERR_STATE = 7,
  RAD_ERR_ADMISSION = 8
};

/* ---- scene description (host side, plain arrays) ------------------------ */

enum { /* Shape factories: "mesh", "sphere", "rectangle", "cube" (build/factory.cpp:266-269) */
  RAD_SHAPE_MESH = 0,
  RAD_SHAPE_SPHERE = 1,
  RAD_SHAPE_RECTANGLE = 2,
  RAD_SHAPE_CUBE = 3
};

/* One scene entity that carries a shape.  Shapes are listed in the order the
 * reference's BFS over entities creates them (build_context.cpp:137-288); the
 * index in this list is the shape index / geomID (embree.cpp:34-36). */
typedef struct rad_shape_desc {
  int32_t type;            /* RAD_SHAPE_* */
  int32_t bsdf;            /* index into rad_scene_desc.bsdfs, -1 = none */
  int32_t light;           /* index into rad_scene_desc.lights, -1 = none */
  int32_t medium_inside;   /* index into media, -1 = none */
  int32_t medium_outside;  /* index into media, -1 = none */
  int32_t flip_normal;     /* rectangle: "is_flip_normal" */
  float to_world[16];      /* entity to_world, children composed as Child*Parent */
  /* mesh: object-space vertex data (world transform applied by the consumer,
   * as Mesh does at construction, shape/mesh.cpp:28-44) */
  uint32_t vertex_count;
  uint32_t triangle_count;
  const float* positions;  /* vertex_count * 3 */
  const float* normals;    /* vertex_count * 3, optional (NULL) */
  const float* uvs;        /* vertex_count * 2, optional (NULL) */
  const uint32_t* indices; /* triangle_count * 3 */
  /* sphere */
  float center[3];
  float radius;
} rad_shape_desc;

enum { /* BSDF factories (build/factory.cpp:274-283) on the hot path */
  RAD_BSDF_DIFFUSE = 0,
  RAD_BSDF_ROUGH_METAL = 1,
  RAD_BSDF_ROUGH_GLASS = 2,
  RAD_BSDF_DISNEY = 3,
  RAD_BSDF_PLASTIC = 4,        /* bsdf/plastic.cpp */
  RAD_BSDF_ROUGH_PLASTIC = 5,  /* bsdf/rough_plastic.cpp */
  RAD_BSDF_PERFECT_MIRROR = 6, /* bsdf/perfect_mirror.cpp */
  RAD_BSDF_PERFECT_GLASS = 7,  /* bsdf/perfect_glass.cpp */
  RAD_BSDF_TWO_SIDE = 8,       /* bsdf/two_side.cpp ("two_side" form): wraps bsdfs[nested] */
  RAD_BSDF_MASK = 9            /* bsdf/mask.cpp: wraps bsdfs[nested] with a constant opacity */
};

enum { RAD_MICROFACET_GGX = 0, RAD_MICROFACET_BECKMANN = 1 };

/* Disney "has value" flags: the reference enables a lobe only when the JSON
 * names the parameter with a non-zero value (bsdf/disney.cpp:459-471). */
enum {
  RAD_DISNEY_HAS_ANISOTROPIC = 1u << 0,
  RAD_DISNEY_HAS_SPEC_TRANS = 1u << 1,
  RAD_DISNEY_HAS_SHEEN = 1u << 2,
  RAD_DISNEY_HAS_SHEEN_TINT = 1u << 3,
  RAD_DISNEY_HAS_FLATNESS = 1u << 4,
  RAD_DISNEY_HAS_SPEC_TINT = 1u << 5,
  RAD_DISNEY_HAS_METALLIC = 1u << 6,
  RAD_DISNEY_HAS_CLEARCOAT = 1u << 7,
  RAD_DISNEY_HAS_ETA = 1u << 8 /* "eta" given (else "specular" is used) */
};

/* BSDF parameters that may be textured (every Texture<T> member of the
 * reference's BSDFs; "alpha" maps to both alpha slots). */
enum {
  RAD_TEX_REFLECTANCE = 0, RAD_TEX_TRANSMITTANCE, RAD_TEX_ETA, RAD_TEX_K, RAD_TEX_ALPHA_U, RAD_TEX_ALPHA_V,
  RAD_TEX_BASE_COLOR, RAD_TEX_ROUGHNESS, RAD_TEX_ANISOTROPIC, RAD_TEX_SPEC_TRANS, RAD_TEX_SHEEN, RAD_TEX_SHEEN_TINT,
  RAD_TEX_FLATNESS, RAD_TEX_SPEC_TINT, RAD_TEX_METALLIC, RAD_TEX_CLEARCOAT, RAD_TEX_CLEARCOAT_GLOSS,
  RAD_TEX_DIFFUSE, RAD_TEX_SPECULAR, RAD_TEX_OPACITY,
  RAD_TEX_SLOTS
};

/* BSDF record.  Defaults are the reference constructors'. */
typedef struct rad_bsdf_desc {
  int32_t type;            /* RAD_BSDF_* */
  int32_t distribution;    /* RAD_MICROFACET_* (rough_metal, rough_glass) */
  int32_t sample_visible;  /* visible-normal sampling (default 1) */
  uint32_t disney_flags;   /* RAD_DISNEY_HAS_* */
  float reflectance[3];    /* diffuse: 0.5; rough_metal / rough_glass: 1 */
  float transmittance[3];  /* rough_glass: 1 */
  float eta[3];            /* rough_metal complex IOR (real) */
  float k[3];              /* rough_metal complex IOR (imaginary) */
  float alpha_u, alpha_v;  /* microfacet roughness (default 0.1) */
  float int_ior, ext_ior;  /* rough_glass (1.5046, 1.000277) */
  /* disney (bsdf/disney.cpp:62-110) */
  float base_color[3];
  float roughness, anisotropic, spec_trans, sheen, sheen_tint, flatness;
  float spec_tint, metallic, clearcoat, clearcoat_gloss;
  float disney_eta;        /* used when RAD_DISNEY_HAS_ETA */
  float specular;          /* used otherwise (default 0.5) */
  float specular_sampling_rate, clearcoat_sampling_rate, diffuse_sampling_rate;
  /* plastic / rough_plastic (plastic.cpp:20-40): int_ior / ext_ior above (1.49, 1.000277),
     rough_plastic's roughness in alpha_u, distribution, sample_visible */
  float diffuse[3];        /* "diffuse" (default 0.5) */
  float specular_reflectance[3]; /* "specular" (default 1) */
  int32_t nonlinear;       /* "no_linear" */
  /* two_side / mask */
  int32_t nested;          /* index of the wrapped BSDF in the scene's bsdf array */
  float opacity;           /* mask "opacity" (default 0.5, clamped to [0, 1]) */
  /* Textured parameters (ConfigNodeReadTexture, build/config_node_ext.h:17-45):
   * tex[RAD_TEX_*] = 1 + index into rad_scene_desc.textures, 0 = the constant above. */
  int32_t tex[RAD_TEX_SLOTS];
} rad_bsdf_desc;

enum {
  RAD_LIGHT_AREA = 0,
  RAD_LIGHT_SKYBOX = 1,
  RAD_LIGHT_POINT = 2,      /* light/point.cpp */
  RAD_LIGHT_PROJECTION = 3  /* light/projection.cpp (constant irradiance) */
};

typedef struct rad_light_desc {
  int32_t type;            /* RAD_LIGHT_* */
  int32_t shape;           /* area: emitting shape index (set by the loader) */
  float radiance[3];       /* area / skybox radiance; point intensity; projection irradiance x scale,
                              or only the scale when the irradiance is a texture */
  float to_world[16];      /* skybox orientation; point: world position in the translation column;
                              projection: the light's frame (looks along +z) */
  float fov;               /* projection field of view in degrees (default 90) */
  int32_t tex;             /* area / skybox / projection: 1 + index of an RGB texture for the radiance
                              (irradiance), 0 = the constant (light/area.cpp:20-41, light/skybox.cpp:24-49,
                              light/projection.cpp:21-26; the projection's aspect is the texture's) */
} rad_light_desc;

enum { RAD_MEDIUM_HOMOGENEOUS = 0, RAD_MEDIUM_HETEROGENEOUS = 1 };
enum { RAD_PHASE_ISOTROPIC = 0, RAD_PHASE_HG = 1 };
enum { RAD_WRAP_CLAMP = 0, RAD_WRAP_REPEAT = 1 };

typedef struct rad_medium_desc {
  int32_t type;            /* RAD_MEDIUM_* */
  int32_t phase;           /* RAD_PHASE_* */
  float g;                 /* hg_phase g, in (-1, 1) */
  float sigma_t[3];        /* constant sigma_t (homogeneous; heterogeneous without grid) */
  float albedo[3];         /* constant single-scatter albedo (default 0.75) */
  float scale;             /* sigma_t scale (default 1) */
  float to_world[16];      /* heterogeneous: entity transform (data lives in [-1,1]^3) */
  int32_t grid_size[3];    /* volume_grid dimensions x, y, z (0 = no grid) */
  int32_t grid_channels;   /* 1 or 3 */
  int32_t grid_wrap;       /* RAD_WRAP_* */
  const float* grid;       /* ((z*Y + y)*X + x)*C floats */
} rad_medium_desc;

/* Textures (texture/bitmap.cpp, texture/chessboard.cpp).  A bitmap holds
 * linear texel values as the reference's image assets deliver them
 * (image_stb.cpp: 8-bit / 255, sRGB -> linear for 3 channels, rows flipped
 * so that texel row 0 is the image's bottom row); texel (x, y) channel c is
 * texels[(y * width + x) * channels + c].  The library builds the mip chain
 * (bitmap.cpp:25-61) itself. */
enum { RAD_TEXTURE_BITMAP = 0, RAD_TEXTURE_CHESSBOARD = 1 };
enum { RAD_FILTER_NEAREST = 0, RAD_FILTER_BILINEAR = 1, RAD_FILTER_TRILINEAR = 2, RAD_FILTER_ANISOTROPIC = 3 };
typedef struct rad_texture_desc {
  int32_t type;            /* RAD_TEXTURE_* */
  int32_t channels;        /* 3 (colour parameters) or 1 (scalar parameters) */
  /* bitmap */
  uint32_t width, height;
  const float* texels;
  int32_t filter;          /* RAD_FILTER_* (default bilinear) */
  int32_t wrap;            /* RAD_WRAP_* (default clamp) */
  int32_t max_level;       /* mip levels: -1 = full chain (trilinear / anisotropic default), 0 = level 0 only */
  float anisotropic_level; /* default 8 for anisotropic, else 0 */
  /* chessboard */
  float to_uv[16];         /* "to_uv" (row-major, default identity) */
  int32_t tex_a, tex_b;    /* 1 + index of a bitmap texture for color_a / color_b, 0 = the constants */
  float color_a[3], color_b[3]; /* defaults 0.4 / 0.2 (mono: channel 0) */
  int32_t remap;           /* "remap" */
} rad_texture_desc;

typedef struct rad_camera_desc {
  float to_world[16];      /* camera-to-world (LookAtLeftHand or explicit) */
  float fov;               /* degrees, horizontal (camera/perspective.cpp:35) */
  float near_clip;         /* default 0.001 */
  float far_clip;          /* default 100 */
  int32_t width, height;
} rad_camera_desc;

typedef struct rad_scene_desc {
  uint32_t shape_count;
  const rad_shape_desc* shapes;
  uint32_t bsdf_count;
  const rad_bsdf_desc* bsdfs;
  uint32_t light_count;
  const rad_light_desc* lights;
  uint32_t medium_count;
  const rad_medium_desc* media;
  int32_t global_medium;   /* index into media, -1 = none */
  rad_camera_desc camera;
  uint32_t texture_count;
  const rad_texture_desc* textures;
} rad_scene_desc;

/* ---- rays and hits --------------------------------------------------------- */

/* 32-byte ray record (two float4): origin, tmin, direction, tmax. */
typedef struct rad_ray {
  float o[3];
  float tmin;
  float d[3];
  float tmax;
} rad_ray;

/* Closest-hit record: geom == 0xffffffff means miss (embree.cpp:96-108). */
typedef struct rad_hit {
  float t, u, v;
  uint32_t prim;
  uint32_t geom;
} rad_hit;

/* ---- render job -------------------------------------------------------------- */

enum { RAD_INTEGRATOR_PATH = 0, RAD_INTEGRATOR_VOL_PATH = 1 };

enum {
  /* Fix the reference's quirks instead of reproducing them (the
   * Frame::SinCosPhi clamp of frame.cpp:48-56 and the first-channel-only
   * Inf/negative sample filter of spectrum.h:30-31).  Off by default. */
  RAD_RENDER_FIX_QUIRKS = 1u << 0,
  /* Also return per-pixel sum of squares (for variance / z-tests). */
  RAD_RENDER_WANT_SUMSQ = 1u << 1,
  /* Time every wavefront kernel with HIP events on the job's stream
   * (rad_job_stats.kernel_ms). */
  RAD_RENDER_PROFILE = 1u << 2,
  /* Shade every path with one kernel instead of sorting the queue by
   * material into specialised kernels (for comparison). */
  RAD_RENDER_UNSORTED = 1u << 3,
  /* path, material-sorted: after every material classification a check
   * verifies that the material queues and the pass-through (suspended) part of
   * the next extension queue hold each extension-queue entry exactly once, in
   * the queue its BSDF type selects; counts go to rad_job_stats.queue_* */
  RAD_RENDER_CHECK_QUEUES = 1u << 4,
  /* with RAD_RENDER_CHECK_QUEUES: count every extension-queue entry twice, so
   * that every entry fails the check (shows that the check can fail) */
  RAD_RENDER_CHECK_SELFTEST = 1u << 5
};

enum { /* rad_job_stats.kernel_ms slots */
  RAD_KERNEL_REGEN = 0,
  RAD_KERNEL_EXTEND = 1,
  RAD_KERNEL_SHADE = 2,
  RAD_KERNEL_SHADOW = 3,
  RAD_KERNEL_GUIDE = 4,    /* guided-vertex gathering + network inference */
  RAD_KERNEL_TRAIN = 5,    /* guide training steps between passes */
  RAD_KERNEL_COUNT = 8
};

typedef struct rad_render_desc {
  int32_t integrator;      /* RAD_INTEGRATOR_* ("path" / "vol_path") */
  int32_t max_depth;       /* renderer.max_depth, -1 = unbounded */
  int32_t rr_depth;        /* renderer.rr_depth (path 3, vol_path 5) */
  uint32_t spp;            /* sampler.sample_count */
  uint32_t seed;           /* sampler.seed */
  uint32_t flags;          /* RAD_RENDER_* */
  uint32_t path_slots;     /* paths in flight (wavefront slots), 0 = chosen by the library */
  int32_t guiding;         /* learned path guiding: 0 off, 1 train online and guide, 2 guide with the scene's current weights */
  uint32_t request_id;     /* RNG stream key for serving (0 for single renders) */
  int32_t width, height;   /* image size override (0 = the scene camera's); fov and pose are the scene's */
  uint32_t blocks_per_cu;        /* persistent grid of the per-path kernels (0 = tuned / default) */
  uint32_t trace_blocks_per_cu;  /* persistent grid of the traversal kernels (0 = tuned / default) */
} rad_render_desc;

typedef struct rad_job_stats {
  double render_ms;        /* device time of the render loop */
  double bvh_build_ms;     /* last BVH build of the scene */
  uint64_t samples;        /* camera samples traced */
  uint64_t segments;       /* extension rays traced (path segments) */
  uint64_t shadow_rays;    /* NEE rays traced */
  uint64_t medium_steps;   /* delta-tracking collisions */
  uint32_t iterations;     /* wavefront iterations */
  uint32_t slots;          /* wavefront capacity */
  double kernel_ms[RAD_KERNEL_COUNT]; /* summed kernel time per stage (RAD_RENDER_PROFILE) */
  uint32_t kernel_launches[RAD_KERNEL_COUNT];
  uint32_t passes;         /* sample passes (result-plane batches) */
  uint32_t guide_steps;    /* training steps the scene's guide has taken in total */
  uint64_t guide_records;  /* training records collected by this render */
  float guide_loss;        /* mean loss of the last training step */
  float guide_alpha;       /* probability of sampling the guide in the final pass */
  uint32_t blocks_per_cu, trace_blocks_per_cu; /* launch configuration used */
  uint64_t suspended;      /* extension traversals cut off at the per-launch node-visit budget and resumed
                              by the next wavefront iteration (results do not depend on it) */
  uint64_t queue_checked;  /* RAD_RENDER_CHECK_QUEUES: extension-queue entries checked */
  uint64_t queue_errors;   /* RAD_RENDER_CHECK_QUEUES: slots missing, duplicated or in the wrong queue */
} rad_job_stats;

/* Launch-configuration autotuner.  Times short renders of `probe` (its spp is
 * the per-trial sample count) over path-slot counts and persistent grid sizes
 * by coordinate descent, stores the fastest configuration in the scene (used
 * by later renders whose desc leaves those fields 0) and reports every trial.
 * Images do not depend on these settings. */
typedef struct rad_tune_trial {
  uint32_t path_slots, blocks_per_cu, trace_blocks_per_cu;
  float ms;
} rad_tune_trial;
typedef struct rad_tune_report {
  uint32_t trials;
  rad_tune_trial trial[24];
  uint32_t path_slots, blocks_per_cu, trace_blocks_per_cu; /* chosen */
  float best_ms;
} rad_tune_report;

enum { RAD_BVH_DEVICE = 0, RAD_BVH_HOST = 1 };
typedef struct rad_bvh_opts {
  int32_t builder;         /* RAD_BVH_DEVICE: GPU binned SAH (3-axis centroid bins, exact sweeps for small
                              subtrees); RAD_BVH_HOST: CPU binned SAH */
  int32_t max_leaf_size;   /* 0 = default (4) */
} rad_bvh_opts;         /* opts = NULL selects the device builder */

/* Structural check of the built BVH (node / primitive buffers read back). */
typedef struct rad_bvh_report {
  uint64_t node_count, primitive_count;
  uint32_t depth, max_leaf;
  uint32_t structure_errors;         /* out-of-range or repeated node references, unreachable nodes */
  uint32_t unreferenced_or_repeated; /* primitives not referenced exactly once by the leaves */
  uint32_t child_box_violations;     /* child boxes not inside the parent's slot box */
  uint32_t leaf_box_violations;      /* triangles / spheres not inside their leaf box */
  float sah_cost;                    /* traversal 1, intersection 1, relative to the root box area */
  int32_t builder;
  uint32_t quantized_box_violations; /* children of the traversal's 64-byte nodes whose dequantised box
                                        does not contain the full-precision one */
} rad_bvh_report;
typedef struct rad_scene_info {
  uint64_t primitive_count;
  uint64_t node_count;
  uint64_t device_bytes;
  double bvh_build_ms;
  float world_min[3];
  float world_max[3];
  uint32_t bvh_depth;
  float sah_cost;
  uint32_t bvh_stack_need; /* worst-case traversal stack entries of the tree (at most 64) */
  uint32_t reserved;
} rad_scene_info;

typedef struct rad_ctx rad_ctx;
typedef struct rad_scene rad_scene;
typedef struct rad_job rad_job;

/* ---- entry points -------------------------------------------------------------- */

/* Thread-local description of the last failure on this thread. */
const char* rad_hip_last_error(void);

/* sizeof() of a struct of this header by name ("rad_shape_desc", ...), 0 if
 * unknown: lets bindings in other languages verify their layouts. */
size_t rad_abi_sizeof(const char* struct_name);

/* Open a device (RadCoreInit analogue).  Fails with RAD_ERR_HIP when no GPU. */
rad_status rad_hip_init(int32_t device_ordinal, rad_ctx** out);
void rad_hip_shutdown(rad_ctx* ctx);

/* Upload a scene (BuildContext::Build minus the accel, build_context.cpp:88-316). */
rad_status rad_hip_scene_create(rad_ctx* ctx, const rad_scene_desc* desc, rad_scene** out);
/* Build the acceleration structure (replaces the Embree ctor, embree.cpp:18-43). */
rad_status rad_hip_bvh_build(rad_scene* scene, const rad_bvh_opts* opts);
rad_status rad_hip_scene_info(const rad_scene* scene, rad_scene_info* info);
rad_status rad_hip_autotune(rad_scene* scene, const rad_render_desc* probe, rad_tune_report* report);
rad_status rad_hip_bvh_check(rad_scene* scene, rad_bvh_report* report);
void rad_hip_scene_destroy(rad_scene* scene);

/* Closest hit for n rays (Embree::RayIntersectPreliminary, embree.cpp:87-111).
 * Host pointers; the call copies in, traces and copies out. */
rad_status rad_hip_intersect_closest(rad_scene* scene, const rad_ray* rays, uint32_t n, rad_hit* hits);
/* Occlusion for n rays (Embree::RayIntersect, embree.cpp:79-85): 1 = occluded. */
rad_status rad_hip_intersect_any(rad_scene* scene, const rad_ray* rays, uint32_t n, uint8_t* occluded);
/* rad_hip_intersect_closest without any device allocation: rays and hits are
 * staged in a static device array of the library (262,144 rays per call at
 * most, 13 MiB per GPU, held until the process ends); more rays fail with
 * RAD_ERR_INVALID_ARG.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Every call on one GPU shares that
 * GPU's array, so calls that overlap on one GPU race. */
rad_status rad_hip_intersect_closest_noalloc(rad_scene* scene, const rad_ray* rays, uint32_t n, rad_hit* hits);
/* Device-pointer variants on a caller stream (hipStream_t as void*, NULL = default). */
rad_status rad_hip_intersect_closest_device(rad_scene* scene, const rad_ray* d_rays, uint32_t n,
                                            rad_hit* d_hits, void* stream);
rad_status rad_hip_intersect_any_device(rad_scene* scene, const rad_ray* d_rays, uint32_t n,
                                        uint8_t* d_occluded, void* stream);

/* Asynchronous render (SampleRenderer::Start, sample_renderer.cpp:60-122). */
rad_status rad_hip_render_start(rad_scene* scene, const rad_render_desc* desc, rad_job** out);
/* Progress in camera samples (AllTaskCount / CompletedTaskCount). */
rad_status rad_hip_job_poll(rad_job* job, uint64_t* done, uint64_t* total, int32_t* complete);
/* Cooperative cancel (Renderer::Stop, sample_renderer.cpp:30-32). */
rad_status rad_hip_job_cancel(rad_job* job);
/* Join and read back: mean radiance W*H*3 (x fastest), optional per-pixel sum
 * of squared sample values (same layout), stats (any may be NULL). */
rad_status rad_hip_job_wait(rad_job* job, float* fb_rgb, float* fb_sumsq, rad_job_stats* stats);
void rad_hip_job_destroy(rad_job* job);

/* ---- JSON scene loading (BuildContext::SetFromJson + Build, host side) -------- */

typedef struct rad_loaded_scene rad_loaded_scene;
/* Parse a reference scene JSON (and its OBJ / mitsuba .vol assets) into a
 * scene description and a render description. */
rad_status rad_scene_load_json(const char* path, rad_loaded_scene** out);
const rad_scene_desc* rad_loaded_scene_desc(const rad_loaded_scene* s);
const rad_render_desc* rad_loaded_scene_render(const rad_loaded_scene* s);
/* renderer.type of the JSON ("path", "vol_path", ...) and renderer.device ("hip" default). */
const char* rad_loaded_scene_renderer_type(const rad_loaded_scene* s);
const char* rad_loaded_scene_device(const rad_loaded_scene* s);
void rad_loaded_scene_free(rad_loaded_scene* s);

/* Write an RGB float image as OpenEXR (Renderer::SaveResult, sample_renderer.cpp:34-46). */
rad_status rad_write_exr(const char* path, const float* rgb, int32_t width, int32_t height);

/* ---- per-function device probes (same device code the integrators run) ------- */

/* Evaluate BSDF records on the device: in = n x {wi[3], wo[3]} (shading space),
 * out = n x {f[3], pdf}.  quirks = 1 reproduces the reference's SinCosPhi clamp. */
rad_status rad_hip_bsdf_eval(const rad_bsdf_desc* bsdf, uint32_t quirks, const float* in, uint32_t n, float* out);
/* Sample: in = n x {wi[3], lobe_xi, xi[2]}, out = n x {wo[3], pdf, f[3], eta}. */
rad_status rad_hip_bsdf_sample(const rad_bsdf_desc* bsdf, uint32_t quirks, const float* in, uint32_t n, float* out);
/* Heterogeneous grid lookup (volume/grid.cpp:41-75): in = n x {p[3]} in [0,1]^3, out = n x 3. */
rad_status rad_hip_grid_eval(const rad_medium_desc* medium, const float* in, uint32_t n, float* out);
/* Texture lookup (Texture<T>::Eval, texture.h:57-62, with bitmap.cpp / chessboard.cpp):
 * textures[index] of an array of count, in = n x {uv[2], duvdx[2], duvdy[2]}, out = n x 3
 * (mono textures fill channel 0). */
rad_status rad_hip_texture_eval(const rad_texture_desc* textures, uint32_t count, uint32_t index, const float* in,
                                uint32_t n, float* out);
/* Texture coordinates at the first hit of the camera ray through each screen
 * position with their screen-space derivatives (perspective.cpp:84-99 and
 * SurfaceInteraction::BSDF(const RayDifferential&), interaction.cpp:83-111):
 * in = n x {sx, sy} pixels, out = n x {uv[2], duvdx[2], duvdy[2]} (zeros on a miss). */
rad_status rad_hip_camera_footprint(rad_scene* scene, const float* in, uint32_t n, float* out);


/* ---- render serving ---------------------------------------------------------------
 * No counterpart in the reference (its CLI renders one scene per process,
 * rad.offline.cli/main.cpp:8-71).  A server admits requests (a resident scene +
 * a render description) in arrival order while their HBM footprint fits the
 * budget, runs admitted requests concurrently on their own streams, reuses the
 * arenas of finished requests, and reports p50 / p99 latency and throughput.
 * A request's image depends only on its scene and description. */
typedef struct rad_server rad_server;
enum { RAD_REQ_QUEUED = 0, RAD_REQ_RUNNING = 1, RAD_REQ_DONE = 2, RAD_REQ_FAILED = 3 };
typedef struct rad_server_opts {
  uint64_t hbm_budget;     /* bytes of job arenas in flight or pooled (0 = 32 GiB) */
  uint32_t max_inflight;   /* concurrent requests (0 = 8) */
  uint32_t reserved;
} rad_server_opts;
typedef struct rad_request_times {
  double queue_ms, run_ms, latency_ms; /* submit -> admit -> done */
} rad_request_times;
typedef struct rad_server_stats {
  uint64_t submitted, completed, failed, queued, running;
  uint64_t samples;        /* camera samples of completed requests */
  double wall_ms;          /* first submit to last completion */
  double samples_per_s;
  double p50_ms, p99_ms, mean_ms; /* request latency (submit -> done) */
  uint64_t held_bytes, peak_held_bytes, budget_bytes;
} rad_server_stats;
rad_status rad_hip_server_create(rad_ctx* ctx, const rad_server_opts* opts, rad_server** out);
/* Queue a render; request_id 0 in desc is replaced by the server's id (the RNG stream key). */
rad_status rad_hip_server_submit(rad_server* server, rad_scene* scene, const rad_render_desc* desc, uint64_t* request);
rad_status rad_hip_server_poll(rad_server* server, uint64_t request, int32_t* state);
/* Block until the request finishes; copy its image (W x H x 3) and release it. */
rad_status rad_hip_server_wait(rad_server* server, uint64_t request, float* fb_rgb, float* fb_sumsq,
                               rad_job_stats* stats, rad_request_times* times);
rad_status rad_hip_server_stats(rad_server* server, rad_server_stats* out);
void rad_hip_server_destroy(rad_server* server);

/* ---- learned path guiding ------------------------------------------------------
 * No counterpart in the reference (its path integrator samples the BSDF only,
 * path.cpp:192-210).  A 5-layer MLP (80-64-64-64-64-48, FP16 MFMA, FP32
 * accumulation) maps a surface vertex to an 8-lobe vMF mixture that is
 * combined with BSDF sampling by one-sample MIS when rad_render_desc.guiding
 * is set.  The scene keeps the trained weights between renders.  The probes
 * below run the same kernels the renderer runs. */
enum { RAD_GUIDE_PARAMS = 20784, RAD_GUIDE_RAW = 48, RAD_GUIDE_FEATS = 9 };
typedef struct rad_guide_record {
  float p[3], n[3], wi[3]; /* vertex position, geometric normal, outgoing (-ray) direction */
  float dir[3];            /* sampled incident direction */
  float thr[3], lbase[3];  /* training bookkeeping (unused by the probes) */
  float w;                 /* sample weight F / p */
  uint32_t prev;
} rad_guide_record;
/* raw = n x 48 network outputs for feats = n x 9 {p, n, wi}; bounds = {lo[3], hi[3]}. */
rad_status rad_hip_guide_forward(const float* params, const float* bounds, const float* feats, uint32_t n, float* raw);
/* Mean loss (1/n) sum_j min(w_j * wnorm, 64) * -log q(dir_j) and its gradient (RAD_GUIDE_PARAMS floats). */
rad_status rad_hip_guide_grad(const float* params, const float* bounds, const rad_guide_record* recs, uint32_t n,
                              float wnorm, float* grad, float* loss);
/* Mixture of one raw row (40 used values): pdf of n directions / n samples {dir[3], pdf} from xi = n x 3. */
rad_status rad_hip_guide_pdf(const float* raw, const float* dirs, uint32_t n, float* pdf);
rad_status rad_hip_guide_sample(const float* raw, const float* xi, uint32_t n, float* out);
/* The scene's guide weights: read (with the number of training steps taken) or replace (NULL = re-initialise). */
rad_status rad_hip_scene_guide_get(rad_scene* scene, float* params, uint32_t* steps);
rad_status rad_hip_scene_guide_set(rad_scene* scene, const float* params);

#ifdef __cplusplus
}
#endif

#endif /* RAD_HIP_H_ */
