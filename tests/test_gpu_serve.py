"""Render serving: continuous admission by HBM footprint, concurrent requests
of different scenes / resolutions / spp, and isolation — a request renders
the same image alone and in a mix of others (bit-identical here: each sample
has its own result entry and the RNG key includes the request id)."""
import time

import numpy as np
import pytest

from radiantiumplusplus_amd import _ffi as F
from radiantiumplusplus_amd import hip, scenes, serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def resident(hip_ctx):
    builders = {
        "many_balls": scenes.many_balls(seed=0, width=96, height=96, spp=8, n_spheres=24),
        "volume_box": scenes.volume_box(width=64, height=64, res=24, spp=8),
        "room": scenes.guiding_room(width=80, height=45, spp=8),
    }
    res = {k: hip.Scene.from_builder(hip_ctx, b) for k, b in builders.items()}
    return builders, res


def desc(b, spp, seed, request_id, w=0, h=0, guiding=0):
    d = b.render_desc(spp=spp, seed=seed, request_id=request_id)
    d.width, d.height, d.guiding = w, h, guiding
    return d


def test_same_image_alone_and_in_the_mix(hip_ctx, resident):
    builders, res = resident
    srv = serve.Server(hip_ctx, max_inflight=6)
    probe = desc(builders["many_balls"], 64, 3, 777, 128, 96)
    alone = srv.wait(srv.submit(res["many_balls"], probe)).image
    rng = np.random.default_rng(0)
    ids = []
    for k in range(10):
        key = ["many_balls", "volume_box", "room"][k % 3]
        w, h = [(64, 64), (96, 72), (160, 90)][rng.integers(3)]
        ids.append(srv.submit(res[key], desc(builders[key], int(rng.choice([4, 16, 32])), k, 1000 + k, w, h)))
        if k == 5:
            mix_id = srv.submit(res["many_balls"], probe)
    out = [srv.wait(i) for i in ids]
    mixed = srv.wait(mix_id).image
    st = srv.stats()
    srv.close()
    assert all(np.isfinite(r.image).all() for r in out)
    assert np.array_equal(alone, mixed)
    assert st.completed == 12 and st.failed == 0
    assert st.p99_ms >= st.p50_ms > 0


def test_admission_respects_the_hbm_budget(hip_ctx, resident):
    builders, res = resident
    b = builders["room"]
    one = desc(b, 64, 1, 5)
    # the footprint of this request shape, from a server that admits anything
    probe = serve.Server(hip_ctx)
    probe.wait(probe.submit(res["room"], one))
    fp = probe.stats().peak_held_bytes
    probe.close()
    budget = int(2.5 * fp)  # room for two such requests at a time
    srv = serve.Server(hip_ctx, hbm_budget=budget, max_inflight=16)
    ids = [srv.submit(res["room"], desc(b, 64, k, 10 + k)) for k in range(8)]
    seen_running = 0
    deadline = time.time() + 60
    while time.time() < deadline:
        st = srv.stats()
        seen_running = max(seen_running, st.running)
        assert st.held_bytes <= budget
        if st.completed + st.failed == 8:
            break
        time.sleep(0.002)
    for i in ids:
        srv.wait(i)
    st = srv.stats()
    srv.close()
    assert st.peak_held_bytes <= budget
    assert st.completed == 8
    assert seen_running <= 2


def test_size_override_matches_a_scene_of_that_size(hip_ctx, resident):
    builders, res = resident
    b = scenes.many_balls(seed=0, width=96, height=96, spp=8, n_spheres=24)
    b2 = scenes.many_balls(seed=0, width=120, height=80, spp=8, n_spheres=24)
    s2 = hip.Scene.from_builder(hip_ctx, b2)
    a = res["many_balls"].render(desc(b, 16, 2, 9, 120, 80))[0]
    c = s2.render(desc(b2, 16, 2, 9))[0]
    assert a.shape == (80, 120, 3)
    assert np.array_equal(a, c)


def test_oversized_request_is_refused(hip_ctx, resident):
    builders, res = resident
    srv = serve.Server(hip_ctx, hbm_budget=1 << 20)
    with pytest.raises(hip.RadError) as e:
        srv.submit(res["room"], desc(builders["room"], 16, 1, 1))
    assert e.value.status == 3  # RAD_ERR_OUT_OF_MEMORY
    srv.close()
