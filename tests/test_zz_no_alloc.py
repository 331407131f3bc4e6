"""rad_hip_intersect_closest_noalloc (static staging, no device allocation)
against rad_hip_intersect_closest on the same rays: bit-identical hits on two
calls in a row, the second with fewer rays, through the C ABI; a batch above
the staging capacity fails with RAD_ERR_INVALID_ARG."""
import ctypes as C

import numpy as np
import pytest

from radiantiumplusplus_amd import hip, scenes

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 18


def noalloc(scene, rays):
    L = hip.lib()
    fn = L.rad_hip_intersect_closest_noalloc
    fn.restype = C.c_int32
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    rays = np.ascontiguousarray(rays, hip.RAY_DTYPE)
    hits = np.zeros(len(rays), hip.HIT_DTYPE)
    st = fn(scene.h, rays.ctypes.data, len(rays), hits.ctypes.data)
    return st, hits


def ray_batch(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, hip.RAY_DTYPE)
    ext = hi - lo
    r["o"] = rng.uniform(lo - 0.1 * ext, hi + 0.1 * ext, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    r["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["tmax"] = np.where(rng.random(n) < 0.2, rng.uniform(0.1, 5.0, n), 3.0e38).astype(np.float32)
    return r


def test_noalloc_matches_allocating_entry_point():
    ctx = hip.Context(0)
    try:
        b = scenes.many_balls(seed=3, width=16, height=16, n_spheres=48)
        s = hip.Scene.from_builder(ctx, b)
        info = s.info()
        lo, hi = np.array(info.world_min[:]), np.array(info.world_max[:])
        for n, seed in ((MAX_RAYS, 1), (1000, 2)):  # the second call is smaller
            rays = ray_batch(lo, hi, n, seed)
            st, got = noalloc(s, rays)
            assert st == 0, hip.lib().rad_hip_last_error()
            ref = s.closest(rays)
            assert got.tobytes() == ref.tobytes()
            assert (ref["geom"] != 0xFFFFFFFF).mean() > 0.1
        st, _ = noalloc(s, ray_batch(lo, hi, 0, 3))
        assert st == 0
        st, _ = noalloc(s, ray_batch(lo, hi, MAX_RAYS + 1, 4))
        assert st == 1  # RAD_ERR_INVALID_ARG
        assert b"staging capacity" in hip.lib().rad_hip_last_error()
    finally:
        ctx.close()
