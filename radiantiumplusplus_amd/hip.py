"""Python front end over the C ABI of libradhip.so (include/rad_hip.h).

Mirrors the reference's renderer plugin contract: a scene is built once
(BuildContext::Build), a render job is started asynchronously (Renderer::Start),
polled (AllTaskCount / CompletedTaskCount / IsComplete), cancelled (Stop) and
joined (Wait).  There is no CPU fallback: without the native library or a
device every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

from . import _ffi as F

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libradhip.so")
_lib = None
_lock = threading.Lock()

RAY_DTYPE = np.dtype([("o", np.float32, 3), ("tmin", np.float32), ("d", np.float32, 3), ("tmax", np.float32)])
HIT_DTYPE = np.dtype([("t", np.float32), ("u", np.float32), ("v", np.float32), ("prim", np.uint32), ("geom", np.uint32)])
MISS = 0xFFFFFFFF


class RadError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"rad status {status}: {msg}")
        self.status = status


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                raise RuntimeError(f"native library missing: {_LIB_PATH} (run __graft_entry__.build())")
            L = C.CDLL(_LIB_PATH)
            vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int32
            sig = {
                "rad_hip_last_error": (C.c_char_p, []),
                "rad_hip_init": (i32, [i32, C.POINTER(vp)]),
                "rad_hip_shutdown": (None, [vp]),
                "rad_hip_scene_create": (i32, [vp, C.POINTER(F.SceneDesc), C.POINTER(vp)]),
                "rad_hip_bvh_build": (i32, [vp, C.POINTER(F.BvhOpts)]),
                "rad_hip_scene_info": (i32, [vp, C.POINTER(F.SceneInfo)]),
                "rad_hip_scene_destroy": (None, [vp]),
                "rad_hip_intersect_closest": (i32, [vp, vp, u32, vp]),
                "rad_hip_intersect_any": (i32, [vp, vp, u32, vp]),
                "rad_hip_intersect_closest_device": (i32, [vp, vp, u32, vp, vp]),
                "rad_hip_intersect_any_device": (i32, [vp, vp, u32, vp, vp]),
                "rad_hip_render_start": (i32, [vp, C.POINTER(F.RenderDesc), C.POINTER(vp)]),
                "rad_hip_job_poll": (i32, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(i32)]),
                "rad_hip_job_cancel": (i32, [vp]),
                "rad_hip_job_wait": (i32, [vp, vp, vp, C.POINTER(F.JobStats)]),
                "rad_hip_job_destroy": (None, [vp]),
                "rad_hip_bsdf_eval": (i32, [C.POINTER(F.BsdfDesc), u32, vp, u32, vp]),
                "rad_hip_bsdf_sample": (i32, [C.POINTER(F.BsdfDesc), u32, vp, u32, vp]),
                "rad_hip_grid_eval": (i32, [C.POINTER(F.MediumDesc), vp, u32, vp]),
                "rad_hip_texture_eval": (i32, [C.POINTER(F.TextureDesc), u32, u32, vp, u32, vp]),
                "rad_hip_camera_footprint": (i32, [vp, vp, u32, vp]),
                "rad_scene_load_json": (i32, [C.c_char_p, C.POINTER(vp)]),
                "rad_loaded_scene_desc": (C.POINTER(F.SceneDesc), [vp]),
                "rad_loaded_scene_render": (C.POINTER(F.RenderDesc), [vp]),
                "rad_loaded_scene_renderer_type": (C.c_char_p, [vp]),
                "rad_loaded_scene_device": (C.c_char_p, [vp]),
                "rad_loaded_scene_free": (None, [vp]),
                "rad_write_exr": (i32, [C.c_char_p, vp, i32, i32]),
                "rad_hip_guide_forward": (i32, [vp, vp, vp, u32, vp]),
                "rad_hip_guide_grad": (i32, [vp, vp, vp, u32, C.c_float, vp, C.POINTER(C.c_float)]),
                "rad_hip_guide_pdf": (i32, [vp, vp, u32, vp]),
                "rad_hip_guide_sample": (i32, [vp, vp, u32, vp]),
                "rad_hip_scene_guide_get": (i32, [vp, vp, C.POINTER(u32)]),
                "rad_hip_scene_guide_set": (i32, [vp, vp]),
                "rad_hip_bvh_check": (i32, [vp, C.POINTER(F.BvhReport)]),
                "rad_hip_autotune": (i32, [vp, C.POINTER(F.RenderDesc), C.POINTER(F.TuneReport)]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
        return _lib


def check(status):
    if status != F.RAD_OK:
        raise RadError(status, lib().rad_hip_last_error().decode(errors="replace"))


class Context:
    """An opened device (RadCoreInit analogue)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().rad_hip_init(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().rad_hip_shutdown(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A scene resident in HBM plus its acceleration structure."""

    def __init__(self, ctx: Context, desc: F.SceneDesc, keepalive=None, builder=F.RAD_BVH_DEVICE, max_leaf=0,
                 build_bvh=True):
        self.ctx = ctx
        self.desc = desc
        self._keep = keepalive
        h = C.c_void_p()
        check(lib().rad_hip_scene_create(ctx.h, C.byref(desc), C.byref(h)))
        self.h = h
        if build_bvh:
            self.build_bvh(builder, max_leaf)

    @classmethod
    def from_builder(cls, ctx, builder_obj, **kw):
        desc = builder_obj.to_desc()
        return cls(ctx, desc, keepalive=builder_obj, **kw)

    def build_bvh(self, builder=F.RAD_BVH_DEVICE, max_leaf=0):
        opts = F.BvhOpts(builder, max_leaf)
        check(lib().rad_hip_bvh_build(self.h, C.byref(opts)))

    def info(self):
        i = F.SceneInfo()
        check(lib().rad_hip_scene_info(self.h, C.byref(i)))
        return i

    def closest(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        check(lib().rad_hip_intersect_closest(self.h, rays.ctypes.data, len(rays), hits.ctypes.data))
        return hits

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(len(rays), np.uint8)
        check(lib().rad_hip_intersect_any(self.h, rays.ctypes.data, len(rays), out.ctypes.data))
        return out

    def camera_footprint(self, screen):
        """uv and d(uv)/dx, d(uv)/dy at the first hit of camera rays through screen positions (n x 2)."""
        pts = np.ascontiguousarray(screen, np.float32)
        out = np.zeros((len(pts), 6), np.float32)
        check(lib().rad_hip_camera_footprint(self.h, pts.ctypes.data, len(pts), out.ctypes.data))
        return out

    def autotune(self, probe: F.RenderDesc):
        """Time short renders over launch configurations; the scene keeps the fastest."""
        r = F.TuneReport()
        check(lib().rad_hip_autotune(self.h, C.byref(probe), C.byref(r)))
        return r

    def bvh_check(self):
        """Structural report of the built BVH (rad_hip_bvh_check)."""
        r = F.BvhReport()
        check(lib().rad_hip_bvh_check(self.h, C.byref(r)))
        return r

    def guide_params(self):
        """The scene's guide weights (flat FP32) and training steps taken."""
        p = np.zeros(F.RAD_GUIDE_PARAMS, np.float32)
        steps = C.c_uint32()
        check(lib().rad_hip_scene_guide_get(self.h, p.ctypes.data, C.byref(steps)))
        return p, steps.value

    def set_guide_params(self, params=None):
        """Replace the guide weights (None = re-initialise)."""
        if params is None:
            check(lib().rad_hip_scene_guide_set(self.h, None))
        else:
            p = np.ascontiguousarray(params, np.float32)
            check(lib().rad_hip_scene_guide_set(self.h, p.ctypes.data))

    def start(self, render_desc: F.RenderDesc):
        return Job(self, render_desc)

    def render(self, render_desc: F.RenderDesc, want_sumsq=False):
        job = self.start(render_desc)
        try:
            return job.wait(want_sumsq=want_sumsq)
        finally:
            job.close()

    def close(self):
        if self.h:
            lib().rad_hip_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job:
    """An asynchronous render (Renderer::Start / Wait / Stop)."""

    def __init__(self, scene: Scene, desc: F.RenderDesc):
        self.scene = scene
        self.desc = desc
        h = C.c_void_p()
        check(lib().rad_hip_render_start(scene.h, C.byref(desc), C.byref(h)))
        self.h = h
        cam = scene.desc.camera
        self.width = desc.width if desc.width > 0 else cam.width
        self.height = desc.height if desc.height > 0 else cam.height

    def poll(self):
        d, t, c = C.c_uint64(), C.c_uint64(), C.c_int32()
        check(lib().rad_hip_job_poll(self.h, C.byref(d), C.byref(t), C.byref(c)))
        return d.value, t.value, bool(c.value)

    def cancel(self):
        check(lib().rad_hip_job_cancel(self.h))

    def wait(self, want_sumsq=False):
        fb = np.zeros((self.height, self.width, 3), np.float32)
        sq = np.zeros((self.height, self.width, 3), np.float32) if want_sumsq else None
        st = F.JobStats()
        check(lib().rad_hip_job_wait(self.h, fb.ctypes.data, sq.ctypes.data if sq is not None else None, C.byref(st)))
        return fb, sq, st

    def close(self):
        if self.h:
            lib().rad_hip_job_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bsdf_eval(desc, wi, wo, quirks=True):
    n = len(wi)
    inp = np.ascontiguousarray(np.concatenate([wi, wo], 1), np.float32)
    out = np.zeros((n, 4), np.float32)
    check(lib().rad_hip_bsdf_eval(C.byref(desc), 1 if quirks else 0, inp.ctypes.data, n, out.ctypes.data))
    return out


def bsdf_sample(desc, wi, lobe, xi, quirks=True):
    n = len(wi)
    inp = np.ascontiguousarray(np.concatenate([wi, lobe.reshape(-1, 1), xi], 1), np.float32)
    out = np.zeros((n, 8), np.float32)
    check(lib().rad_hip_bsdf_sample(C.byref(desc), 1 if quirks else 0, inp.ctypes.data, n, out.ctypes.data))
    return out


def texture_eval(textures, index, lookups):
    """textures: ctypes array of TextureDesc; lookups: n x {uv[2], duvdx[2], duvdy[2]} -> n x 3."""
    q = np.ascontiguousarray(lookups, np.float32)
    out = np.zeros((len(q), 3), np.float32)
    check(lib().rad_hip_texture_eval(textures, len(textures), index, q.ctypes.data, len(q), out.ctypes.data))
    return out


def grid_eval(medium_desc, pts):
    pts = np.ascontiguousarray(pts, np.float32)
    out = np.zeros((len(pts), 3), np.float32)
    check(lib().rad_hip_grid_eval(C.byref(medium_desc), pts.ctypes.data, len(pts), out.ctypes.data))
    return out
