"""`path` on the GPU against the oracle's restatement of Path::Li
(renderer/path.cpp:131-227): per-tile z-test at 4 sigma and relMSE < 1e-3 on
tile means, plus closed-form scenes and the render-job contract."""
import time

import numpy as np
import pytest

import _oracle as O
import stats
from radiantiumplusplus_amd import _ffi as F
from radiantiumplusplus_amd import hip, scenes

pytestmark = pytest.mark.gpu


def gpu_render(ctx, b, spp, seed=1, flags=0):
    s = hip.Scene.from_builder(ctx, b)
    rd = b.render_desc(spp=spp, seed=seed, flags=flags | F.RAD_RENDER_WANT_SUMSQ)
    fb, sq, st = s.render(rd, want_sumsq=True)
    return fb, sq, st


def oracle_render(b, spp, seed=99, quirks=True):
    o = O.OracleScene(b, quirks=quirks)
    fb, sq, _ = o.render(spp=spp, seed=seed)
    return fb, sq


def assert_same_up_to_rounding(a, u):
    """Material-sorted and single-kernel shading run separate compilations of
    the same arithmetic, whose float contractions may differ by an ulp."""
    d = np.abs(a - u)
    assert d.max() <= 1e-6 * max(1.0, float(np.abs(a).max())), d.max()
    assert np.count_nonzero(d) <= max(8, a.size // 1000)


def assert_match(g, o, spp_g, spp_o, tile=8, zmax=4.0, relmax=1e-3):
    z, rel = stats.compare(g[0], g[1], spp_g, o[0], o[1], spp_o, tile=tile)
    print(f"max |z| {np.abs(z).max():.2f}  relMSE {rel:.2e}")
    assert np.abs(z).max() < zmax, f"tile z-test failed: max |z| = {np.abs(z).max():.2f}"
    assert rel < relmax, f"relMSE {rel:.3e}"


def test_many_balls_matches_oracle(hip_ctx):
    b = scenes.many_balls(seed=0, width=64, height=64, spp=16)
    g = gpu_render(hip_ctx, b, 4096)
    o = oracle_render(b, 1024)
    assert np.isfinite(g[0]).all()
    assert_match(g, o, 4096, 1024)


def test_glossy_materials_match_oracle(hip_ctx):
    """GGX conductor, rough dielectric and Disney with the reference's
    visible-normal quirk reproduced."""
    rng = np.random.default_rng(5)
    b = scenes.SceneBuilder()
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(6, 6, 6)), bsdf=b.diffuse(0.6))
    mats = [b.rough_metal(alpha=0.2), b.rough_glass(alpha=0.15, int_ior=1.5, ext_ior=1.0),
            b.disney(base_color=(0.7, 0.3, 0.2), roughness=0.3, metallic=0.5, clearcoat=0.6, clearcoat_gloss=0.7),
            b.rough_metal(alpha=0.4, distribution="beckmann")]
    for i, m in enumerate(mats):
        p, n, uv, idx = scenes.create_sphere(0.8, 24)
        b.mesh(p, idx, n, uv, to_world=scenes.trs(translate=(-3 + 2 * i, 0.8, 0)), bsdf=m)
    b.sphere(center=(0, 0.6, -2.0), radius=0.6, bsdf=b.disney(base_color=(0.2, 0.6, 0.8), spec_trans=0.8, roughness=0.2))
    light = b.rectangle(to_world=scenes.trs(translate=(0, 6, 0), axis=(1, 0, 0), angle=90, scale=(1.5, 1.5, 1.5)))
    b.area_light(light, 8.0)
    b.skybox((0.2, 0.25, 0.3))
    b.set_camera(64, 48, 50, origin=(0, 3, -8), target=(0, 0.7, 0), far=1000)
    b.set_render("path", spp=16, max_depth=12, rr_depth=3, seed=3)
    g = gpu_render(hip_ctx, b, 4096)
    o = oracle_render(b, 1024)
    assert_match(g, o, 4096, 1024, relmax=2e-3)


def test_fixed_quirks_match_oracle(hip_ctx):
    b = scenes.many_balls(seed=2, width=48, height=48, spp=16, n_spheres=16)
    for i in range(1, len(b.bsdfs), 2):
        b.bsdfs[i] = {"type": "rough_metal", "alpha": 0.25, "reflectance": [0.9, 0.8, 0.6], "eta": [0.2, 0.9, 1.1],
                      "k": [3.9, 2.4, 2.1], "distribution": "ggx", "sample_visible": True}
    g = gpu_render(hip_ctx, b, 2048, flags=F.RAD_RENDER_FIX_QUIRKS)
    o = oracle_render(b, 1024, quirks=False)
    assert_match(g, o, 2048, 1024, relmax=2e-3)


def test_furnace_closed_form(hip_ctx):
    rho = 0.6
    b = scenes.SceneBuilder()
    b.sphere(center=(0, 0, 0), radius=1.0, bsdf=b.diffuse(rho))
    b.skybox(1.0)
    b.set_camera(32, 32, 40, origin=(0, 0, -4), target=(0, 0, 0), far=100)
    b.set_render("path", spp=256, max_depth=-1, rr_depth=1000, seed=7)
    fb, sq, _ = gpu_render(hip_ctx, b, 1024)
    assert np.allclose(fb[0, 0], 1.0, rtol=1e-6)
    c = fb[14:18, 14:18].mean(axis=(0, 1))
    assert np.allclose(c, rho, atol=0.01)


def test_job_progress_and_cancel(hip_ctx):
    b = scenes.many_balls(seed=0, width=128, height=128, spp=16)
    s = hip.Scene.from_builder(hip_ctx, b)
    # Tens of seconds of work, so the cancel always lands while it runs.
    spp = 1 << 20
    job = s.start(b.render_desc(spp=spp, seed=1))
    time.sleep(0.2)
    done, total, complete = job.poll()
    assert total == 128 * 128 * spp
    assert not complete and done < total
    job.cancel()
    with pytest.raises(hip.RadError) as e:
        job.wait()
    assert e.value.status == 6  # RAD_ERR_CANCELLED
    job.close()
    fb, _, st = s.render(b.render_desc(spp=8, seed=1))
    assert st.samples == 128 * 128 * 8
    assert st.segments > st.samples


def test_deterministic_per_seed(hip_ctx):
    b = scenes.many_balls(seed=1, width=32, height=32, spp=8, n_spheres=16)
    s = hip.Scene.from_builder(hip_ctx, b)
    a = s.render(b.render_desc(spp=8, seed=3))[0]
    c = s.render(b.render_desc(spp=8, seed=3))[0]
    d = s.render(b.render_desc(spp=8, seed=4))[0]
    assert np.array_equal(a, c)
    assert not np.array_equal(a, d)


def test_json_scene_end_to_end(hip_ctx, tmp_path):
    """The JSON front end and the in-memory description render identically."""
    import ctypes as C

    b = scenes.many_balls(seed=4, width=32, height=32, spp=8, n_spheres=12, slices=16)
    path = b.write_json(str(tmp_path))
    h = C.c_void_p()
    hip.check(hip.lib().rad_scene_load_json(path.encode(), C.byref(h)))
    try:
        d = hip.lib().rad_loaded_scene_desc(h).contents
        rd = hip.lib().rad_loaded_scene_render(h).contents
        s1 = hip.Scene(hip_ctx, d)
        fb1 = s1.render(rd)[0]
        s2 = hip.Scene.from_builder(hip_ctx, b)
        fb2 = s2.render(b.render_desc())[0]
        assert np.array_equal(fb1, fb2)
    finally:
        hip.lib().rad_loaded_scene_free(h)


def test_autotune_and_launch_independence(hip_ctx):
    """The autotuner picks a launch configuration the scene keeps; images do
    not depend on slots, grid sizes or material sorting."""
    b = scenes.many_balls(seed=0, width=256, height=256, spp=16, n_spheres=24)
    s = hip.Scene.from_builder(hip_ctx, b)
    rep = s.autotune(b.render_desc(spp=16, seed=1))
    print("trials", [(t.path_slots, t.blocks_per_cu, t.trace_blocks_per_cu, round(t.ms, 2))
                     for t in rep.trial[:rep.trials]], "chosen", rep.path_slots, rep.blocks_per_cu,
          rep.trace_blocks_per_cu)
    assert rep.trials >= 8
    assert rep.path_slots in (1 << 20, 2 << 20, 4 << 20, 8 << 20)
    a, _, st = s.render(b.render_desc(spp=16, seed=5))
    assert (st.blocks_per_cu, st.trace_blocks_per_cu) == (rep.blocks_per_cu, rep.trace_blocks_per_cu)
    d = b.render_desc(spp=16, seed=5)
    d.path_slots, d.blocks_per_cu, d.trace_blocks_per_cu = 4096, 2, 2
    c, _, st2 = s.render(d)
    assert st2.slots == 4096
    assert np.array_equal(a, c)
    u = s.render(b.render_desc(spp=16, seed=5, flags=F.RAD_RENDER_UNSORTED))[0]
    assert_same_up_to_rounding(a, u)


def remaining_types_scene():
    """plastic, rough_plastic, perfect_mirror, perfect_glass and mask spheres on
    a diffuse floor, a two_side panel seen from its back, lit by an area light,
    a point light and a projection light."""
    b = scenes.SceneBuilder()
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(6, 6, 6)), bsdf=b.diffuse(0.6))
    mats = [b.plastic(diffuse=(0.7, 0.2, 0.2)), b.rough_plastic(alpha=0.3, diffuse=(0.2, 0.6, 0.3)),
            b.perfect_mirror(reflectance=(0.9, 0.9, 0.8), eta=(0.2, 0.9, 1.1), k=(3.9, 2.4, 2.1)),
            b.perfect_glass(int_ior=1.5, ext_ior=1.0), b.mask(b.diffuse((0.3, 0.3, 0.8)), opacity=0.5)]
    for i, m in enumerate(mats):
        b.sphere(center=(-3.2 + 1.6 * i, 0.7, 0), radius=0.7, bsdf=m)
    # panel facing away from the camera (+z): one-sided diffuse would be black from here
    b.rectangle(to_world=scenes.trs(translate=(0, 1.5, 2.5), scale=(2.5, 1.5, 1)), bsdf=b.two_side(b.diffuse(0.7)))
    light = b.rectangle(to_world=scenes.trs(translate=(0, 5, -1), axis=(1, 0, 0), angle=90, scale=(1, 1, 1)))
    b.area_light(light, 6.0)
    b.point_light((2.5, 3.0, -2.0), (12.0, 10.0, 8.0))
    tw = np.array([[1, 0, 0, -2.0], [0, 0, -1, 4.0], [0, 1, 0, -1.0], [0, 0, 0, 1]], np.float32)  # looks down
    b.projection_light(tw, irradiance=(0.8, 0.9, 1.0), scale=2.0, fov=35.0)
    b.set_camera(64, 40, 55, origin=(0, 3, -8), target=(0, 0.8, 0), far=1000)
    b.set_render("path", spp=16, max_depth=10, rr_depth=3, seed=3)
    return b


@pytest.mark.parametrize("fix", [False, True])
def test_remaining_bsdfs_and_lights_match_oracle(hip_ctx, fix):
    b = remaining_types_scene()
    flags = F.RAD_RENDER_FIX_QUIRKS if fix else 0
    g = gpu_render(hip_ctx, b, 4096, flags=flags)
    o = oracle_render(b, 1024, quirks=not fix)
    assert np.isfinite(g[0]).all()
    assert_match(g, o, 4096, 1024, relmax=2e-3)


def test_remaining_types_sorted_equals_unsorted(hip_ctx):
    """The generic material queue (every type without a queue of its own)
    gives the same image as the single shading kernel."""
    b = remaining_types_scene()
    s = hip.Scene.from_builder(hip_ctx, b)
    a = s.render(b.render_desc(spp=8, seed=5))[0]
    u = s.render(b.render_desc(spp=8, seed=5, flags=F.RAD_RENDER_UNSORTED))[0]
    assert_same_up_to_rounding(a, u)


@pytest.mark.parametrize("budget", [1, 3, 24])
def test_suspended_traversals_leave_images_unchanged(hip_ctx, monkeypatch, budget):
    """Extension traversals cut off at a per-launch node-visit budget resume
    in the next iteration exactly where they stopped: the image and the ray
    count equal those of uncut traversals, sorted or not, guided or not.
    Budget 1 also overflows the resume pool, whose refused rays finish in
    their own launch."""
    b = scenes.many_balls(seed=0, width=128, height=128, spp=16, n_spheres=200)
    s = hip.Scene.from_builder(hip_ctx, b)
    s.set_guide_params(None)
    for guiding in (0, 2):
        def desc(flags=0):
            d = b.render_desc(spp=8, seed=5, flags=flags)
            d.guiding = guiding
            return d
        for flags in (0, F.RAD_RENDER_UNSORTED):
            monkeypatch.setenv("RAD_VISIT_BUDGET", "0")
            ref, _, st0 = s.render(desc(flags))
            assert st0.suspended == 0
            monkeypatch.setenv("RAD_VISIT_BUDGET", str(budget))
            img, _, st = s.render(desc(flags))
            print(f"guiding {guiding} flags {flags} budget {budget}: suspended {st.suspended} "
                  f"iterations {st0.iterations} -> {st.iterations}")
            assert st.suspended > 0
            assert st.segments == st0.segments
            assert np.array_equal(img, ref)


@pytest.mark.parametrize("budget", ["0", "2"])
def test_material_queues_are_a_permutation(hip_ctx, monkeypatch, budget):
    """After every material classification the five material queues and the
    pass-through (suspended) entries of the next extension queue hold each
    extension-queue slot exactly once, each in the queue its BSDF type
    selects (a device-side tally; budget 2 makes suspended traversals pass
    through).  The check leaves the image unchanged, and the self-test mode,
    which counts every entry twice, shows that it can fail."""
    monkeypatch.setenv("RAD_VISIT_BUDGET", budget)
    b = remaining_types_scene()
    pos, _, _, idx = scenes.create_sphere(0.6, 48)  # a 2 k-triangle mesh: traversals deep enough to suspend
    b.mesh(pos, idx, to_world=scenes.trs(translate=(2.2, 0.6, -1.5)), bsdf=b.rough_metal(alpha=0.2))
    s = hip.Scene.from_builder(hip_ctx, b)
    ref, _, st0 = s.render(b.render_desc(spp=8, seed=5))
    img, _, st = s.render(b.render_desc(spp=8, seed=5, flags=F.RAD_RENDER_CHECK_QUEUES))
    print(f"budget {budget}: checked {st.queue_checked} errors {st.queue_errors} "
          f"segments {st.segments} suspended {st.suspended}")
    assert st.queue_errors == 0
    assert st.queue_checked == st.segments + st.suspended and st.segments > 0
    assert (st.suspended > 0) == (budget != "0")
    assert np.array_equal(img, ref)
    _, _, bad = s.render(b.render_desc(spp=8, seed=5, flags=F.RAD_RENDER_CHECK_QUEUES | F.RAD_RENDER_CHECK_SELFTEST))
    assert bad.queue_errors == bad.queue_checked > 0
    # the unsorted kernel has no material queues: nothing is checked
    _, _, un = s.render(b.render_desc(spp=8, seed=5, flags=F.RAD_RENDER_CHECK_QUEUES | F.RAD_RENDER_UNSORTED))
    assert un.queue_checked == 0
