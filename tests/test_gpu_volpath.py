"""`vol_path` on the GPU against the oracle's restatement of VolPath::Li
(renderer/vol_path.cpp:27-349): delta tracking through heterogeneous and
homogeneous media, pass-through medium boundaries, a global medium, the
transmittance loop of the light sampler, plus a closed-form absorber."""
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
    return s.render(rd, want_sumsq=True)


def assert_match(b, spp_g, spp_o, ctx, tile=8, zmax=4.0, relmax=1e-3):
    fb, sq, st = gpu_render(ctx, b, spp_g)
    o = O.OracleScene(b)
    ofb, osq, _ = o.render(spp=spp_o, seed=77)
    assert np.isfinite(fb).all()
    z, rel = stats.compare(fb, sq, spp_g, ofb, osq, spp_o, tile=tile)
    print(f"max |z| {np.abs(z).max():.2f}  relMSE {rel:.2e}  medium steps {st.medium_steps}")
    assert np.abs(z).max() < zmax, f"tile z-test failed: max |z| = {np.abs(z).max():.2f}"
    assert rel < relmax, f"relMSE {rel:.3e}"
    return st


def test_heterogeneous_box_matches_oracle(hip_ctx):
    b = scenes.volume_box(width=48, height=48, res=32, spp=64)
    st = assert_match(b, 2048, 512, hip_ctx)
    assert st.medium_steps > 0


def test_homogeneous_sphere_in_fog_matches_oracle(hip_ctx):
    """Scattering HG medium inside a BSDF-less sphere, an isotropic global
    fog, a glossy floor (BSDF NEE through the fog) and a skybox."""
    b = scenes.SceneBuilder()
    fog = b.homogeneous(sigma_t=0.08, albedo=0.9)
    ball = b.homogeneous(sigma_t=(1.5, 2.0, 3.0), albedo=(0.9, 0.8, 0.7), g=0.5)
    b.global_medium = fog
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(5, 5, 5)), bsdf=b.rough_metal(alpha=0.3))
    b.sphere(center=(0, 1.0, 0), radius=1.0, medium_inside=ball, medium_outside=fog)
    light = b.rectangle(to_world=scenes.trs(translate=(0, 4, 0), axis=(1, 0, 0), angle=90, scale=(1, 1, 1)))
    b.area_light(light, 10.0)
    b.skybox((0.1, 0.12, 0.15))
    b.set_camera(40, 32, 45, origin=(0, 2, -6), target=(0, 1, 0), far=1000)
    b.set_render("vol_path", spp=64, max_depth=24, rr_depth=5, seed=4)
    assert_match(b, 2048, 512, hip_ctx)


def test_absorber_closed_form(hip_ctx):
    """A purely absorbing slab (albedo 0, sigma_t 1, thickness 1) in front of
    a unit emitter: delta tracking escapes with probability exp(-1).  The
    1-degree field of view keeps every camera ray within 1e-4 of normal
    incidence (a wider one lengthens off-axis paths through the slab)."""
    b = scenes.SceneBuilder()
    slab = b.homogeneous(sigma_t=1.0, albedo=0.0)
    b.cube(to_world=scenes.trs(scale=(20, 20, 0.5)), medium_inside=slab)
    em = b.rectangle(to_world=scenes.trs(translate=(0, 0, 3), axis=(0, 1, 0), angle=180, scale=(30, 30, 1)))
    b.area_light(em, 1.0)
    b.set_camera(16, 16, 1, origin=(0, 0, -5), target=(0, 0, 0), far=1000)
    b.set_render("vol_path", spp=64, max_depth=-1, rr_depth=1000, seed=2)
    fb, sq, _ = gpu_render(hip_ctx, b, 8192)
    m = fb.mean(axis=(0, 1))
    n = 16 * 16 * 8192
    p = np.exp(-1.0)
    se = np.sqrt(p * (1 - p) / n)
    assert np.all(np.abs(m - p) < 5 * se), (m, p, se)


def test_without_media_matches_path(hip_ctx):
    """With no medium anywhere vol_path reduces to path."""
    b = scenes.many_balls(seed=3, width=48, height=48, spp=16, n_spheres=12)
    b.renderer["rr_depth"] = 5
    s = hip.Scene.from_builder(hip_ctx, b)
    rd = b.render_desc(spp=2048, seed=3, flags=F.RAD_RENDER_WANT_SUMSQ)
    fb_p, sq_p, _ = s.render(rd, want_sumsq=True)
    rd.integrator = F.RAD_INTEGRATOR_VOL_PATH
    rd.seed = 9
    fb_v, sq_v, st = s.render(rd, want_sumsq=True)
    assert st.medium_steps == 0
    z, rel = stats.compare(fb_v, sq_v, 2048, fb_p, sq_p, 2048, tile=8)
    print(f"max |z| {np.abs(z).max():.2f}  relMSE {rel:.2e}")
    assert np.abs(z).max() < 4.0 and rel < 1e-3


def test_volpath_deterministic(hip_ctx):
    b = scenes.volume_box(width=24, height=24, res=16, spp=8)
    s = hip.Scene.from_builder(hip_ctx, b)
    a = s.render(b.render_desc(spp=8, seed=5))[0]
    c = s.render(b.render_desc(spp=8, seed=5))[0]
    assert np.array_equal(a, c)


def test_material_sorted_shading_matches_single_kernel(hip_ctx):
    """vol_path shades each material queue with its own kernel (medium events
    in whichever queue the ray's surface hit put the path); the image equals
    the single generic kernel's up to float contraction.  The scene mixes a
    medium with diffuse, metal, glass and Disney surfaces and a mask."""
    b = scenes.SceneBuilder()
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(4, 4, 4)), bsdf=b.diffuse(0.6))
    mats = [b.rough_metal(alpha=0.3), b.rough_glass(alpha=0.2, int_ior=1.5, ext_ior=1.0), b.disney(),
            b.mask(b.diffuse((0.3, 0.3, 0.8)), opacity=0.5)]
    for i, m in enumerate(mats):
        b.sphere(center=(-1.8 + 1.2 * i, 0.5, 0.8), radius=0.5, bsdf=m)
    b.global_medium = b.homogeneous(sigma_t=0.35, albedo=0.85)
    b.area_light(b.rectangle(to_world=scenes.trs(translate=(0, 3.5, 0), axis=(1, 0, 0), angle=90)), 6.0)
    b.set_camera(32, 24, 45, origin=(0, 1.5, -5), target=(0, 0.5, 0))
    b.set_render("vol_path", spp=8, max_depth=8, rr_depth=4, seed=3)
    s = hip.Scene.from_builder(hip_ctx, b)
    a = s.render(b.render_desc(spp=8, seed=5))[0]
    u = s.render(b.render_desc(spp=8, seed=5, flags=F.RAD_RENDER_UNSORTED))[0]
    d = np.abs(a - u)
    assert d.max() <= 1e-6 * max(1.0, float(np.abs(a).max())), d.max()
    assert np.count_nonzero(d) <= max(8, a.size // 1000)
    assert a.max() > 0
