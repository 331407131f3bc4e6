"""ctypes mirrors of the plain-C structs declared in include/rad_hip.h.

The layouts here must match the header field for field; tests/test_abi.py
checks the sizes against the values the compiled library reports.
"""
import ctypes as C

RAD_OK = 0

RAD_SHAPE_MESH, RAD_SHAPE_SPHERE, RAD_SHAPE_RECTANGLE, RAD_SHAPE_CUBE = 0, 1, 2, 3
RAD_BSDF_DIFFUSE, RAD_BSDF_ROUGH_METAL, RAD_BSDF_ROUGH_GLASS, RAD_BSDF_DISNEY = 0, 1, 2, 3
RAD_BSDF_PLASTIC, RAD_BSDF_ROUGH_PLASTIC, RAD_BSDF_PERFECT_MIRROR, RAD_BSDF_PERFECT_GLASS = 4, 5, 6, 7
RAD_BSDF_TWO_SIDE, RAD_BSDF_MASK = 8, 9
RAD_MICROFACET_GGX, RAD_MICROFACET_BECKMANN = 0, 1
RAD_LIGHT_AREA, RAD_LIGHT_SKYBOX, RAD_LIGHT_POINT, RAD_LIGHT_PROJECTION = 0, 1, 2, 3
RAD_MEDIUM_HOMOGENEOUS, RAD_MEDIUM_HETEROGENEOUS = 0, 1
RAD_PHASE_ISOTROPIC, RAD_PHASE_HG = 0, 1
RAD_WRAP_CLAMP, RAD_WRAP_REPEAT = 0, 1
RAD_INTEGRATOR_PATH, RAD_INTEGRATOR_VOL_PATH = 0, 1
RAD_RENDER_FIX_QUIRKS = 1 << 0
RAD_RENDER_WANT_SUMSQ = 1 << 1

DISNEY_FLAGS = {
    "anisotropic": 1 << 0,
    "spec_trans": 1 << 1,
    "sheen": 1 << 2,
    "sheen_tint": 1 << 3,
    "flatness": 1 << 4,
    "spec_tint": 1 << 5,
    "metallic": 1 << 6,
    "clearcoat": 1 << 7,
    "eta": 1 << 8,
}

F3 = C.c_float * 3
F16 = C.c_float * 16
I3 = C.c_int32 * 3
FP = C.POINTER(C.c_float)
UP = C.POINTER(C.c_uint32)


class ShapeDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("bsdf", C.c_int32),
        ("light", C.c_int32),
        ("medium_inside", C.c_int32),
        ("medium_outside", C.c_int32),
        ("flip_normal", C.c_int32),
        ("to_world", F16),
        ("vertex_count", C.c_uint32),
        ("triangle_count", C.c_uint32),
        ("positions", FP),
        ("normals", FP),
        ("uvs", FP),
        ("indices", UP),
        ("center", F3),
        ("radius", C.c_float),
    ]


class BsdfDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("distribution", C.c_int32),
        ("sample_visible", C.c_int32),
        ("disney_flags", C.c_uint32),
        ("reflectance", F3),
        ("transmittance", F3),
        ("eta", F3),
        ("k", F3),
        ("alpha_u", C.c_float),
        ("alpha_v", C.c_float),
        ("int_ior", C.c_float),
        ("ext_ior", C.c_float),
        ("base_color", F3),
        ("roughness", C.c_float),
        ("anisotropic", C.c_float),
        ("spec_trans", C.c_float),
        ("sheen", C.c_float),
        ("sheen_tint", C.c_float),
        ("flatness", C.c_float),
        ("spec_tint", C.c_float),
        ("metallic", C.c_float),
        ("clearcoat", C.c_float),
        ("clearcoat_gloss", C.c_float),
        ("disney_eta", C.c_float),
        ("specular", C.c_float),
        ("specular_sampling_rate", C.c_float),
        ("clearcoat_sampling_rate", C.c_float),
        ("diffuse_sampling_rate", C.c_float),
        ("diffuse", F3),
        ("specular_reflectance", F3),
        ("nonlinear", C.c_int32),
        ("nested", C.c_int32),
        ("opacity", C.c_float),
        ("tex", C.c_int32 * 20),
    ]


# textured BSDF parameter slots (rad_bsdf_desc.tex)
(RAD_TEX_REFLECTANCE, RAD_TEX_TRANSMITTANCE, RAD_TEX_ETA, RAD_TEX_K, RAD_TEX_ALPHA_U, RAD_TEX_ALPHA_V,
 RAD_TEX_BASE_COLOR, RAD_TEX_ROUGHNESS, RAD_TEX_ANISOTROPIC, RAD_TEX_SPEC_TRANS, RAD_TEX_SHEEN, RAD_TEX_SHEEN_TINT,
 RAD_TEX_FLATNESS, RAD_TEX_SPEC_TINT, RAD_TEX_METALLIC, RAD_TEX_CLEARCOAT, RAD_TEX_CLEARCOAT_GLOSS,
 RAD_TEX_DIFFUSE, RAD_TEX_SPECULAR, RAD_TEX_OPACITY) = range(20)
RAD_TEX_SLOTS = 20
RAD_TEXTURE_BITMAP, RAD_TEXTURE_CHESSBOARD = 0, 1
RAD_FILTER_NEAREST, RAD_FILTER_BILINEAR, RAD_FILTER_TRILINEAR, RAD_FILTER_ANISOTROPIC = 0, 1, 2, 3


class TextureDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("channels", C.c_int32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("texels", C.POINTER(C.c_float)),
        ("filter", C.c_int32),
        ("wrap", C.c_int32),
        ("max_level", C.c_int32),
        ("anisotropic_level", C.c_float),
        ("to_uv", F16),
        ("tex_a", C.c_int32),
        ("tex_b", C.c_int32),
        ("color_a", F3),
        ("color_b", F3),
        ("remap", C.c_int32),
    ]


class LightDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("shape", C.c_int32),
        ("radiance", F3),
        ("to_world", F16),
        ("fov", C.c_float),
        ("tex", C.c_int32),
    ]


class MediumDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("phase", C.c_int32),
        ("g", C.c_float),
        ("sigma_t", F3),
        ("albedo", F3),
        ("scale", C.c_float),
        ("to_world", F16),
        ("grid_size", I3),
        ("grid_channels", C.c_int32),
        ("grid_wrap", C.c_int32),
        ("grid", FP),
    ]


class CameraDesc(C.Structure):
    _fields_ = [
        ("to_world", F16),
        ("fov", C.c_float),
        ("near_clip", C.c_float),
        ("far_clip", C.c_float),
        ("width", C.c_int32),
        ("height", C.c_int32),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("shape_count", C.c_uint32),
        ("shapes", C.POINTER(ShapeDesc)),
        ("bsdf_count", C.c_uint32),
        ("bsdfs", C.POINTER(BsdfDesc)),
        ("light_count", C.c_uint32),
        ("lights", C.POINTER(LightDesc)),
        ("medium_count", C.c_uint32),
        ("media", C.POINTER(MediumDesc)),
        ("global_medium", C.c_int32),
        ("camera", CameraDesc),
        ("texture_count", C.c_uint32),
        ("textures", C.POINTER(TextureDesc)),
    ]


class Ray(C.Structure):
    _fields_ = [("o", F3), ("tmin", C.c_float), ("d", F3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("prim", C.c_uint32), ("geom", C.c_uint32)]


class RenderDesc(C.Structure):
    _fields_ = [
        ("integrator", C.c_int32),
        ("max_depth", C.c_int32),
        ("rr_depth", C.c_int32),
        ("spp", C.c_uint32),
        ("seed", C.c_uint32),
        ("flags", C.c_uint32),
        ("path_slots", C.c_uint32),
        ("guiding", C.c_int32),
        ("request_id", C.c_uint32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("blocks_per_cu", C.c_uint32),
        ("trace_blocks_per_cu", C.c_uint32),
    ]


class JobStats(C.Structure):
    _fields_ = [
        ("render_ms", C.c_double),
        ("bvh_build_ms", C.c_double),
        ("samples", C.c_uint64),
        ("segments", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("medium_steps", C.c_uint64),
        ("iterations", C.c_uint32),
        ("slots", C.c_uint32),
        ("kernel_ms", C.c_double * 8),
        ("kernel_launches", C.c_uint32 * 8),
        ("passes", C.c_uint32),
        ("guide_steps", C.c_uint32),
        ("guide_records", C.c_uint64),
        ("guide_loss", C.c_float),
        ("guide_alpha", C.c_float),
        ("blocks_per_cu", C.c_uint32),
        ("trace_blocks_per_cu", C.c_uint32),
        ("suspended", C.c_uint64),
        ("queue_checked", C.c_uint64),
        ("queue_errors", C.c_uint64),
    ]


RAD_RENDER_PROFILE = 1 << 2
RAD_RENDER_UNSORTED = 1 << 3
RAD_RENDER_CHECK_QUEUES = 1 << 4
RAD_RENDER_CHECK_SELFTEST = 1 << 5
RAD_KERNEL_REGEN, RAD_KERNEL_EXTEND, RAD_KERNEL_SHADE, RAD_KERNEL_SHADOW = 0, 1, 2, 3
RAD_KERNEL_GUIDE, RAD_KERNEL_TRAIN = 4, 5
RAD_BVH_DEVICE, RAD_BVH_HOST = 0, 1
RAD_GUIDE_PARAMS, RAD_GUIDE_RAW, RAD_GUIDE_FEATS = 20784, 48, 9


class GuideRecord(C.Structure):
    _fields_ = [
        ("p", C.c_float * 3),
        ("n", C.c_float * 3),
        ("wi", C.c_float * 3),
        ("dir", C.c_float * 3),
        ("thr", C.c_float * 3),
        ("lbase", C.c_float * 3),
        ("w", C.c_float),
        ("prev", C.c_uint32),
    ]



class BvhOpts(C.Structure):
    _fields_ = [("builder", C.c_int32), ("max_leaf_size", C.c_int32)]


class SceneInfo(C.Structure):
    _fields_ = [
        ("primitive_count", C.c_uint64),
        ("node_count", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("bvh_build_ms", C.c_double),
        ("world_min", F3),
        ("world_max", F3),
        ("bvh_depth", C.c_uint32),
        ("sah_cost", C.c_float),
        ("bvh_stack_need", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


RAY_DTYPE_FIELDS = ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")


RAD_REQ_QUEUED, RAD_REQ_RUNNING, RAD_REQ_DONE, RAD_REQ_FAILED = 0, 1, 2, 3


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_inflight", C.c_uint32), ("reserved", C.c_uint32)]


class RequestTimes(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("run_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("failed", C.c_uint64),
        ("queued", C.c_uint64),
        ("running", C.c_uint64),
        ("samples", C.c_uint64),
        ("wall_ms", C.c_double),
        ("samples_per_s", C.c_double),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("mean_ms", C.c_double),
        ("held_bytes", C.c_uint64),
        ("peak_held_bytes", C.c_uint64),
        ("budget_bytes", C.c_uint64),
    ]


class BvhReport(C.Structure):
    _fields_ = [
        ("node_count", C.c_uint64),
        ("primitive_count", C.c_uint64),
        ("depth", C.c_uint32),
        ("max_leaf", C.c_uint32),
        ("structure_errors", C.c_uint32),
        ("unreferenced_or_repeated", C.c_uint32),
        ("child_box_violations", C.c_uint32),
        ("leaf_box_violations", C.c_uint32),
        ("sah_cost", C.c_float),
        ("builder", C.c_int32),
        ("quantized_box_violations", C.c_uint32),
    ]


class TuneTrial(C.Structure):
    _fields_ = [("path_slots", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("ms", C.c_float)]


class TuneReport(C.Structure):
    _fields_ = [
        ("trials", C.c_uint32),
        ("trial", TuneTrial * 24),
        ("path_slots", C.c_uint32),
        ("blocks_per_cu", C.c_uint32),
        ("trace_blocks_per_cu", C.c_uint32),
        ("best_ms", C.c_float),
    ]


ABI_STRUCTS = {
    "rad_shape_desc": ShapeDesc,
    "rad_bsdf_desc": BsdfDesc,
    "rad_light_desc": LightDesc,
    "rad_medium_desc": MediumDesc,
    "rad_camera_desc": CameraDesc,
    "rad_scene_desc": SceneDesc,
    "rad_ray": Ray,
    "rad_hit": Hit,
    "rad_render_desc": RenderDesc,
    "rad_bvh_opts": BvhOpts,
    "rad_scene_info": SceneInfo,
    "rad_job_stats": JobStats,
    "rad_guide_record": GuideRecord,
    "rad_server_opts": ServerOpts,
    "rad_request_times": RequestTimes,
    "rad_server_stats": ServerStats,
    "rad_bvh_report": BvhReport,
    "rad_tune_report": TuneReport,
}
