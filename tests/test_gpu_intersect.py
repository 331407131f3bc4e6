"""Closest-hit / occlusion parity of the HIP traversal with the oracle's brute
force on seeded ray batches.  (geomID, primID) must match; a mismatch is
allowed only where the two candidates' hit distances tie within 2 ulp."""
import numpy as np
import pytest

import _oracle as O
from radiantiumplusplus_amd import hip, scenes

pytestmark = pytest.mark.gpu


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def ray_batch(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, hip.RAY_DTYPE)
    ext = hi - lo
    r["o"] = rng.uniform(lo - 0.1 * ext, hi + 0.1 * ext, (n, 3)).astype(np.float32)
    r["d"] = unit(rng.normal(size=(n, 3))).astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = np.where(rng.random(n) < 0.2, rng.uniform(0.1, 5.0, n), 3.0e38).astype(np.float32)
    return r


def check_parity(oscene, rays, hg, hb):
    assert hg.shape == hb.shape
    same = (hg["geom"] == hb["geom"]) & (hg["prim"] == hb["prim"])
    # identical arithmetic on both sides: matched hits carry bit-identical t, u, v
    m = same & (hb["geom"] != 0xFFFFFFFF)
    assert np.array_equal(hg["t"][m], hb["t"][m])
    assert np.array_equal(hg["u"][m], hb["u"][m])
    bad = np.nonzero(~same)[0]
    for i in bad:
        assert hg["geom"][i] != 0xFFFFFFFF and hb["geom"][i] != 0xFFFFFFFF, f"hit/miss disagreement at ray {i}"
        tg = oscene.prim_t(rays[i], hg["geom"][i], hg["prim"][i])
        tb = hb["t"][i]
        ulp = np.spacing(np.float32(max(abs(tg), abs(tb))))
        assert abs(tg - tb) <= 2 * ulp, f"ray {i}: t {tg} vs {tb}"
    return len(bad)


@pytest.mark.parametrize("seed", [0, 1])
def test_closest_hit_matches_brute_force(hip_ctx, seed):
    b = scenes.many_balls(seed=seed, n_spheres=64, slices=32, width=16, height=16)
    o = O.OracleScene(b)
    g = hip.Scene.from_builder(hip_ctx, b)
    lo, hi = o.bounds()
    rays = ray_batch(lo, hi, 20000, seed)
    hg = g.closest(rays)
    hb = o.closest(rays, brute=True)
    ties = check_parity(o, rays, hg, hb)
    assert ties <= 10
    assert (hb["geom"] != 0xFFFFFFFF).mean() > 0.2


def test_secondary_rays_from_surfaces(hip_ctx):
    """Rays leaving hit points (the integrator's workload) must agree too."""
    b = scenes.many_balls(seed=3, n_spheres=64, slices=32, width=16, height=16)
    o = O.OracleScene(b)
    g = hip.Scene.from_builder(hip_ctx, b)
    lo, hi = o.bounds()
    prim = ray_batch(lo, hi, 20000, 5)
    h = o.closest(prim, brute=True)
    hit = h["geom"] != 0xFFFFFFFF
    rng = np.random.default_rng(9)
    sec = np.zeros(int(hit.sum()), hip.RAY_DTYPE)
    p = prim["o"][hit] + prim["d"][hit] * h["t"][hit][:, None]
    sec["d"] = unit(rng.normal(size=(len(sec), 3))).astype(np.float32)
    sec["o"] = (p + sec["d"] * 1e-4).astype(np.float32)
    sec["tmax"] = 3.0e38
    hg = g.closest(sec)
    hb = o.closest(sec, brute=True)
    check_parity(o, sec, hg, hb)


def test_occlusion_matches_brute_force(hip_ctx):
    b = scenes.many_balls(seed=4, n_spheres=64, slices=32, width=16, height=16)
    o = O.OracleScene(b)
    g = hip.Scene.from_builder(hip_ctx, b)
    lo, hi = o.bounds()
    rays = ray_batch(lo, hi, 20000, 11)
    occ_g = g.occluded(rays)
    occ_b = o.occluded(rays, brute=True)
    diff = np.nonzero(occ_g != occ_b)[0]
    # an occlusion disagreement can only come from a hit at exactly tmax
    for i in diff:
        h = o.closest(rays[i : i + 1], brute=True)[0]
        assert h["geom"] != 0xFFFFFFFF and abs(h["t"] - rays["tmax"][i]) <= 2 * np.spacing(rays["tmax"][i])
    assert len(diff) <= 5


def test_empty_batch_and_bvh_info(hip_ctx):
    b = scenes.many_balls(seed=0, n_spheres=8, slices=8, width=8, height=8)
    g = hip.Scene.from_builder(hip_ctx, b)
    assert len(g.closest(np.zeros(0, hip.RAY_DTYPE))) == 0
    info = g.info()
    assert info.primitive_count == b.triangle_count + 2
    assert info.node_count > 0 and info.bvh_depth > 0


@pytest.mark.parametrize("pool", [1, 5])
def test_primitives_beyond_the_wave_pool(hip_ctx, monkeypatch, pool):
    """A wave tests up to 64 leaf primitives per node step as one pool; a
    lane whose entries fall beyond it tests them itself.  RAD_TRACE_POOL
    shrinks the pool so that this path carries most primitives: hits must
    still match the brute force, and images must equal the full pool's bit
    for bit (the closest hit is the minimum of (t, primitive) either way)."""
    b = scenes.many_balls(seed=6, n_spheres=64, slices=32, width=24, height=24)
    o = O.OracleScene(b)
    full = hip.Scene.from_builder(hip_ctx, b)
    monkeypatch.setenv("RAD_TRACE_POOL", str(pool))
    small = hip.Scene.from_builder(hip_ctx, b)
    lo, hi = o.bounds()
    rays = ray_batch(lo, hi, 20000, 13)
    check_parity(o, rays, small.closest(rays), o.closest(rays, brute=True))
    occ_b = o.occluded(rays, brute=True)
    assert (small.occluded(rays) != occ_b).sum() <= 5
    d = b.render_desc(spp=8, seed=3)
    img_full = full.render(d)[0]
    img_small = small.render(d)[0]
    assert np.array_equal(img_full, img_small)

