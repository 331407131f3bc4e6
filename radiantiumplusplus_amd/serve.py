"""Render serving over the native server (rad_hip_server_*): requests of
different scenes, resolutions and spp are admitted by HBM footprint as they
arrive and run concurrently; results carry queue / run / latency times and
the server reports p50 / p99 latency beside throughput.

`poisson_requests` draws the seeded request stream of the serving workload:
Poisson arrivals, scene, resolution, spp and guiding drawn per request."""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _ffi as F
from . import hip

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = hip.lib()
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int32
    sig = {
        "rad_hip_server_create": (i32, [vp, C.POINTER(F.ServerOpts), C.POINTER(vp)]),
        "rad_hip_server_submit": (i32, [vp, vp, C.POINTER(F.RenderDesc), C.POINTER(C.c_uint64)]),
        "rad_hip_server_poll": (i32, [vp, C.c_uint64, C.POINTER(i32)]),
        "rad_hip_server_wait": (i32, [vp, C.c_uint64, vp, vp, C.POINTER(F.JobStats), C.POINTER(F.RequestTimes)]),
        "rad_hip_server_stats": (i32, [vp, C.POINTER(F.ServerStats)]),
        "rad_hip_server_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _bound = True


@dataclass
class Result:
    image: np.ndarray
    sumsq: np.ndarray | None
    stats: F.JobStats
    times: F.RequestTimes


class Server:
    def __init__(self, ctx: hip.Context, hbm_budget=0, max_inflight=0):
        _bind()
        self.ctx = ctx
        opts = F.ServerOpts(hbm_budget=hbm_budget, max_inflight=max_inflight)
        h = C.c_void_p()
        hip.check(hip.lib().rad_hip_server_create(ctx.h, C.byref(opts), C.byref(h)))
        self.h = h
        self._shape = {}

    def submit(self, scene: hip.Scene, desc: F.RenderDesc) -> int:
        rid = C.c_uint64()
        hip.check(hip.lib().rad_hip_server_submit(self.h, scene.h, C.byref(desc), C.byref(rid)))
        cam = scene.desc.camera
        w = desc.width if desc.width > 0 else cam.width
        h = desc.height if desc.height > 0 else cam.height
        self._shape[rid.value] = (h, w, bool(desc.flags & F.RAD_RENDER_WANT_SUMSQ))
        return rid.value

    def poll(self, rid: int) -> int:
        st = C.c_int32()
        hip.check(hip.lib().rad_hip_server_poll(self.h, rid, C.byref(st)))
        return st.value

    def wait(self, rid: int) -> Result:
        h, w, want = self._shape.pop(rid)
        fb = np.zeros((h, w, 3), np.float32)
        sq = np.zeros((h, w, 3), np.float32) if want else None
        st, tm = F.JobStats(), F.RequestTimes()
        hip.check(hip.lib().rad_hip_server_wait(self.h, rid, fb.ctypes.data, sq.ctypes.data if sq is not None else None,
                                                C.byref(st), C.byref(tm)))
        return Result(fb, sq, st, tm)

    def stats(self) -> F.ServerStats:
        s = F.ServerStats()
        hip.check(hip.lib().rad_hip_server_stats(self.h, C.byref(s)))
        return s

    def close(self):
        if self.h:
            hip.lib().rad_hip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


RESOLUTIONS = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
SPPS = [4, 16, 64, 256]


@dataclass
class Spec:
    t: float          # arrival time (s) after the stream starts
    scene: str        # key of the resident scene
    width: int
    height: int
    spp: int
    guiding: int


def poisson_requests(n, rate=20.0, seed=4, scenes=("many_balls", "large_mesh", "volume_box"), resolutions=RESOLUTIONS,
                     spps=SPPS, guided=(0, 2)):
    """Seeded request stream: exponential inter-arrival times of mean
    1/rate, scene, resolution, spp and guiding uniform over the given sets
    (guiding only for `path` scenes)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        sc = scenes[rng.integers(len(scenes))]
        w, h = resolutions[rng.integers(len(resolutions))]
        spp = int(spps[rng.integers(len(spps))])
        g = int(guided[rng.integers(len(guided))]) if sc != "volume_box" else 0
        out.append(Spec(t, sc, w, h, spp, g))
    return out


def run_stream(server: Server, scenes: dict, builders: dict, specs, want_sumsq=False):
    """Submit `specs` at their arrival times (real time) and wait for all of
    them; returns the per-request results in submission order."""
    t0 = time.perf_counter()
    ids = []
    for sp in specs:
        delay = sp.t - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        b = builders[sp.scene]
        d = b.render_desc(spp=sp.spp, seed=1, flags=F.RAD_RENDER_WANT_SUMSQ if want_sumsq else 0)
        d.width, d.height = sp.width, sp.height
        d.guiding = sp.guiding
        ids.append(server.submit(scenes[sp.scene], d))
    return [server.wait(i) for i in ids]
