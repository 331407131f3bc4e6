"""Textures on the GPU against the oracle: the texture probe over every filter
and wrap mode (tolerance 2e-5 relative, the float rounding of the blend
weights and of exp / log2), the screen-space footprint of camera rays on
meshes (with and without uvs), spheres and rectangles (rtol 1e-3), textured
renders by the per-tile z-test, and texture-independent invariants (sorted
= unsorted shading, constant textures = constant parameters)."""
import ctypes

import numpy as np
import pytest

import _oracle as O
import stats
from radiantiumplusplus_amd import _ffi as F
from radiantiumplusplus_amd import hip, scenes

pytestmark = pytest.mark.gpu

IDENT16 = [float(x) for x in np.eye(4, dtype=np.float32).reshape(-1)]
KEEP = []


def bitmap(img, filt, wrap, max_level, aniso):
    img = np.ascontiguousarray(img, np.float32)
    KEEP.append(img)
    d = F.TextureDesc()
    d.type = F.RAD_TEXTURE_BITMAP
    d.channels = 3 if img.ndim == 3 else 1
    d.height, d.width = img.shape[:2]
    d.texels = img.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    d.filter, d.wrap, d.max_level, d.anisotropic_level = filt, wrap, max_level, aniso
    d.to_uv[:] = IDENT16
    return d


def lookups(rng, n, scale):
    uv = rng.uniform(-0.4, 1.4, (n, 2))
    dx = rng.normal(0, 1, (n, 2)) * np.exp(rng.uniform(np.log(scale * 1e-3), np.log(scale), (n, 1)))
    dy = rng.normal(0, 1, (n, 2)) * np.exp(rng.uniform(np.log(scale * 1e-3), np.log(scale), (n, 1)))
    return np.concatenate([uv, dx, dy], 1).astype(np.float32)


CASES = [(filt, wrap, shape) for filt in (F.RAD_FILTER_NEAREST, F.RAD_FILTER_BILINEAR, F.RAD_FILTER_TRILINEAR,
                                          F.RAD_FILTER_ANISOTROPIC)
         for wrap in (F.RAD_WRAP_CLAMP, F.RAD_WRAP_REPEAT) for shape in ((64, 32, 3), (16, 16), (12, 10, 3))]


@pytest.mark.parametrize("filt,wrap,shape", CASES)
def test_texture_probe_matches_oracle(hip_ctx, filt, wrap, shape):
    rng = np.random.default_rng(hash((filt, wrap, shape)) % 2**32)
    img = rng.uniform(0, 1, shape).astype(np.float32)
    mip = filt in (F.RAD_FILTER_TRILINEAR, F.RAD_FILTER_ANISOTROPIC)
    arr = (F.TextureDesc * 1)()
    arr[0] = bitmap(img, filt, wrap, -1 if mip else 0, 8.0 if filt == F.RAD_FILTER_ANISOTROPIC else 0.0)
    q = lookups(rng, 4000, 0.05)
    g = hip.texture_eval(arr, 0, q)
    o = O.texture_eval(arr, 0, q)
    err = np.abs(g - o) / np.maximum(np.abs(o), 1e-3)
    print(f"max rel err {err.max():.2e}")
    assert np.isfinite(g).all()
    assert err.max() < 2e-5


def test_chessboard_probe_matches_oracle(hip_ctx):
    rng = np.random.default_rng(11)
    img = rng.uniform(0, 1, (32, 32, 3)).astype(np.float32)
    cb = F.TextureDesc()
    cb.type = F.RAD_TEXTURE_CHESSBOARD
    cb.channels = 3
    cb.to_uv[:] = [float(x) for x in scenes.trs(axis=(0, 0, 1), angle=20, scale=(3, 5, 1)).reshape(-1)]
    cb.color_a[:] = [0.9, 0.1, 0.2]
    cb.tex_b = 2
    cb.remap = 1
    arr = (F.TextureDesc * 2)()
    arr[0] = cb
    arr[1] = bitmap(img, F.RAD_FILTER_ANISOTROPIC, F.RAD_WRAP_REPEAT, -1, 8.0)
    q = lookups(rng, 4000, 0.05)
    g = hip.texture_eval(arr, 0, q)
    o = O.texture_eval(arr, 0, q)
    assert np.abs(g - o).max() < 2e-5


def textured_scene(integrator="path", with_uvs=True):
    """A bitmap floor (trilinear, repeat), a chessboard (nested anisotropic
    bitmap) on a rough-metal sphere with a textured roughness, a textured
    plastic sphere, a Disney mesh with uvs (or barycentric uvs) and a textured
    mask, under an area light."""
    rng = np.random.default_rng(5)
    b = scenes.SceneBuilder()
    img = rng.uniform(0.1, 0.9, (64, 64, 3)).astype(np.float32)
    img[::8] *= 0.3  # stripes that the footprint filters
    mono = rng.uniform(0.05, 0.5, (16, 16)).astype(np.float32)
    floor_tex = b.bitmap(img, filter="trilinear", wrap="repeat")
    cb = b.chessboard(color_a=(0.8, 0.2, 0.2), color_b=b.bitmap(img, filter="anisotropic"),
                      to_uv=scenes.trs(scale=(4, 4, 1)), remap=True)
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(6, 6, 6)), bsdf=b.diffuse(floor_tex))
    b.sphere(center=(-1.2, 1, 0), radius=1, bsdf=b.rough_metal(alpha=b.bitmap(mono, filter="nearest"), reflectance=cb))
    b.sphere(center=(1.3, 0.8, -0.5), radius=0.8, bsdf=b.plastic(diffuse=b.bitmap(img, filter="bilinear")))
    verts, _, sphere_uv, idx = scenes.create_sphere(0.6, 12)
    uvs = sphere_uv if with_uvs else None
    b.mesh(verts, idx, uvs=uvs, to_world=scenes.trs(translate=(0.2, 0.6, 1.5)),
           bsdf=b.disney(base_color=b.bitmap(img, filter="nearest"), roughness=0.4, metallic=b.bitmap(mono)))
    b.rectangle(to_world=scenes.trs(translate=(2.5, 1.2, 1.0), axis=(0, 1, 0), angle=30, scale=(0.7, 0.7, 1)),
                bsdf=b.mask(b.diffuse((0.2, 0.5, 0.8)), opacity=b.bitmap(mono, filter="bilinear")))
    light = b.rectangle(to_world=scenes.trs(translate=(0, 5, -1), axis=(1, 0, 0), angle=90, scale=(1, 1, 1)))
    b.area_light(light, 8.0)
    b.set_camera(64, 40, 55, origin=(0, 3, -8), target=(0, 0.8, 0), far=1000)
    b.set_render(integrator, spp=16, max_depth=8, rr_depth=3, seed=3)
    return b


@pytest.mark.parametrize("with_uvs", [True, False])
def test_camera_footprint_matches_oracle(hip_ctx, with_uvs):
    b = textured_scene(with_uvs=with_uvs)
    s = hip.Scene.from_builder(hip_ctx, b)
    o = O.OracleScene(b)
    rng = np.random.default_rng(2)
    pts = np.stack([rng.uniform(0, 64, 3000), rng.uniform(0, 40, 3000)], 1).astype(np.float32)
    g = s.camera_footprint(pts)
    r = o.camera_footprint(pts)
    hit_r, hit_g = np.any(r != 0, 1), np.any(g != 0, 1)
    assert np.count_nonzero(hit_r != hit_g) <= 3  # grazing rays at silhouettes
    hit = hit_r & hit_g
    g, r = g[hit], r[hit]
    uv_err = np.abs(g[:, :2] - r[:, :2])
    # relative to the footprint's size (single components cancel to ~0)
    scale = np.maximum(np.abs(r[:, 2:]).max(1, keepdims=True), 1e-6)
    d_err = np.abs(g[:, 2:] - r[:, 2:]) / scale
    print(f"uv max err {uv_err.max():.2e}, footprint max rel err {d_err.max():.2e}")
    # uv wraps at the sphere seam (phi = 0 / 2 pi): compare modulo 1 there
    uv_err = np.minimum(uv_err, 1 - uv_err)
    assert uv_err.max() < 2e-4
    assert np.quantile(d_err, 0.999) < 1e-3 and d_err.max() < 1e-2


@pytest.mark.parametrize("integrator", ["path", "vol_path"])
def test_textured_scene_matches_oracle(hip_ctx, integrator):
    b = textured_scene(integrator)
    s = hip.Scene.from_builder(hip_ctx, b)
    rd = b.render_desc(spp=4096, seed=1, flags=F.RAD_RENDER_WANT_SUMSQ)
    g, gsq, _ = s.render(rd, want_sumsq=True)
    o = O.OracleScene(b)
    of, osq, _ = o.render(spp=1024, seed=99)
    z, rel = stats.compare(g, gsq, 4096, of, osq, 1024, tile=8)
    print(f"max |z| {np.abs(z).max():.2f}  relMSE {rel:.2e}")
    assert np.isfinite(g).all()
    assert np.abs(z).max() < 4.0
    assert rel < 2e-3


def test_textured_sorted_equals_unsorted(hip_ctx):
    """Textured BSDFs shade in the generic material kernel; it and the single
    shading kernel are separate compilations of the same code, whose float
    contractions may differ by an ulp in a few pixels."""
    b = textured_scene()
    s = hip.Scene.from_builder(hip_ctx, b)
    a = s.render(b.render_desc(spp=8, seed=5))[0]
    u = s.render(b.render_desc(spp=8, seed=5, flags=F.RAD_RENDER_UNSORTED))[0]
    d = np.abs(a - u)
    assert np.count_nonzero(d) <= 8 and d.max() <= 1e-6 * max(1.0, float(a.max()))


def test_constant_textures_equal_constant_parameters(hip_ctx):
    """1 x 1 nearest textures return their value unchanged, so the render equals
    the untextured scene bit for bit (textured BSDFs take the generic queue,
    which shades the same arithmetic)."""
    import test_texture_oracle as T
    for kind in ("metal", "plastic", "disney"):
        plain, textured = T.textured_pair(kind)
        a = hip.Scene.from_builder(hip_ctx, plain).render(plain.render_desc(spp=16, seed=7))[0]
        t = hip.Scene.from_builder(hip_ctx, textured).render(textured.render_desc(spp=16, seed=7))[0]
        assert np.array_equal(a, t), kind


def test_json_textured_scene_end_to_end(hip_ctx, tmp_path):
    """Textures written as PNG image assets by the builder and read back by the
    JSON front end render identically to the in-memory description."""
    import ctypes as C

    rng = np.random.default_rng(12)
    b = scenes.SceneBuilder()
    img = rng.integers(0, 256, (16, 32, 3)).astype(np.float32) / np.float32(255)
    mono = rng.integers(10, 120, (8, 8)).astype(np.float32) / np.float32(255)
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(6, 6, 6)),
                bsdf=b.diffuse(b.bitmap(img, filter="anisotropic", wrap="repeat")))
    b.sphere(center=(0, 1, 0), radius=1,
             bsdf=b.rough_metal(alpha=b.bitmap(mono, filter="nearest"),
                                reflectance=b.chessboard((0.8, 0.2, 0.2), b.bitmap(img), to_uv=scenes.trs(scale=(4, 4, 1)))))
    light = b.rectangle(to_world=scenes.trs(translate=(0, 5, -1), axis=(1, 0, 0), angle=90))
    b.area_light(light, 8.0)
    b.set_camera(48, 32, 55, origin=(0, 3, -8), target=(0, 0.8, 0), far=1000)
    b.set_render("path", spp=8, max_depth=6, rr_depth=3, seed=3)
    path = b.write_json(str(tmp_path))
    h = C.c_void_p()
    hip.check(hip.lib().rad_scene_load_json(path.encode(), C.byref(h)))
    try:
        d = hip.lib().rad_loaded_scene_desc(h).contents
        rd = hip.lib().rad_loaded_scene_render(h).contents
        fb1 = hip.Scene(hip_ctx, d).render(rd)[0]
        fb2 = hip.Scene.from_builder(hip_ctx, b).render(b.render_desc())[0]
        assert np.array_equal(fb1, fb2)
    finally:
        hip.lib().rad_loaded_scene_free(h)


def emitter_scene(integrator="path"):
    """A textured rectangle emitter (8 x 4 radiance map) and a rotated,
    translated 16 x 8 environment map over a diffuse floor, a rough-metal
    sphere and a textured diffuse sphere."""
    rng = np.random.default_rng(21)
    b = scenes.SceneBuilder()
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(5, 5, 5)), bsdf=b.diffuse(0.6))
    b.sphere(center=(-1.0, 1, 0), radius=1, bsdf=b.rough_metal(alpha=0.3))
    b.sphere(center=(1.4, 0.7, -0.4), radius=0.7, bsdf=b.diffuse(b.bitmap(rng.uniform(0.2, 0.9, (8, 8, 3)))))
    light = b.rectangle(to_world=scenes.trs(translate=(0, 4, -1), axis=(1, 0, 0), angle=80, scale=(1.5, 1, 1)))
    b.area_light(light, b.bitmap(rng.uniform(0, 12, (4, 8, 3)), filter="nearest"))
    env = rng.uniform(0, 1.5, (8, 16, 3)).astype(np.float32)
    env[2:4, 3:6] = 20.0  # a bright region the distribution has to find
    sky_tw = scenes.trs(translate=(0.5, 0.2, -0.3), axis=(0, 1, 0), angle=35)
    b.skybox(b.bitmap(env, filter="bilinear"), to_world=sky_tw)
    b.set_camera(64, 40, 60, origin=(0, 2.5, -7), target=(0, 0.8, 0), far=1000)
    b.set_render(integrator, spp=16, max_depth=6, rr_depth=3, seed=4)
    return b


@pytest.mark.parametrize("integrator,fix", [("path", False), ("path", True), ("vol_path", False)])
def test_textured_emitters_match_oracle(hip_ctx, integrator, fix):
    b = emitter_scene(integrator)
    s = hip.Scene.from_builder(hip_ctx, b)
    flags = F.RAD_RENDER_WANT_SUMSQ | (F.RAD_RENDER_FIX_QUIRKS if fix else 0)
    g, gsq, _ = s.render(b.render_desc(spp=4096, seed=1, flags=flags), want_sumsq=True)
    of, osq, _ = O.OracleScene(b, quirks=not fix).render(spp=1024, seed=99)
    z, rel = stats.compare(g, gsq, 4096, of, osq, 1024, tile=8)
    print(f"max |z| {np.abs(z).max():.2f}  relMSE {rel:.2e}")
    assert np.isfinite(g).all()
    assert np.abs(z).max() < 4.0
    assert rel < 2e-3


def test_textured_environment_furnace_gpu(hip_ctx):
    """Constant 8 x 4 map with quirks off: the sphere shows its albedo."""
    import test_texture_oracle as T
    b = T.env_furnace("bitmap")
    s = hip.Scene.from_builder(hip_ctx, b)
    fb = s.render(b.render_desc(spp=1024, seed=3, flags=F.RAD_RENDER_FIX_QUIRKS))[0]
    assert np.allclose(fb[0, 0], 1.0, rtol=1e-6)
    assert np.allclose(fb[6:10, 6:10].mean(axis=(0, 1)), 0.6, atol=0.01)


def test_json_textured_emitters_end_to_end(hip_ctx, tmp_path):
    import ctypes as C

    rng = np.random.default_rng(22)
    b = scenes.SceneBuilder()
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(5, 5, 5)), bsdf=b.diffuse(0.6))
    light = b.rectangle(to_world=scenes.trs(translate=(0, 4, -1), axis=(1, 0, 0), angle=80))
    b.area_light(light, b.bitmap(rng.integers(0, 256, (4, 8, 3)).astype(np.float32) / np.float32(255)))
    b.skybox(b.bitmap(rng.integers(0, 256, (8, 16, 3)).astype(np.float32) / np.float32(255)))
    b.set_camera(32, 24, 60, origin=(0, 2.5, -7), target=(0, 0.8, 0), far=1000)
    b.set_render("path", spp=8, max_depth=5, rr_depth=3, seed=4)
    path = b.write_json(str(tmp_path))
    h = C.c_void_p()
    hip.check(hip.lib().rad_scene_load_json(path.encode(), C.byref(h)))
    try:
        d = hip.lib().rad_loaded_scene_desc(h).contents
        rd = hip.lib().rad_loaded_scene_render(h).contents
        fb1 = hip.Scene(hip_ctx, d).render(rd)[0]
        fb2 = hip.Scene.from_builder(hip_ctx, b).render(b.render_desc())[0]
        assert np.array_equal(fb1, fb2)
    finally:
        hip.lib().rad_loaded_scene_free(h)


def projection_scene(integrator="path"):
    """Two projection lights with irradiance images over a diffuse floor and
    spheres: an 8 x 4 bilinear bitmap (aspect 2) and a chessboard."""
    rng = np.random.default_rng(23)
    b = scenes.SceneBuilder()
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(5, 5, 5)), bsdf=b.diffuse(0.6))
    b.sphere(center=(-1.0, 0.8, 0.3), radius=0.8, bsdf=b.rough_metal(alpha=0.3))
    b.sphere(center=(1.2, 0.6, -0.5), radius=0.6, bsdf=b.diffuse((0.7, 0.6, 0.5)))
    down = np.array([[1, 0, 0, -0.5], [0, 0, -1, 4.0], [0, 1, 0, 0.2], [0, 0, 0, 1]], np.float32)
    b.projection_light(down, irradiance=b.bitmap(rng.integers(0, 256, (4, 8, 3)).astype(np.float32) / np.float32(255)),
                       scale=(6.0, 5.0, 4.0), fov=50.0)
    slant = np.array([[1, 0, 0, 1.5], [0, -0.6, -0.8, 3.0], [0, 0.8, -0.6, 2.0], [0, 0, 0, 1]], np.float32)
    b.projection_light(slant, irradiance=b.chessboard((3.0, 2.0, 1.0), (0.2, 0.5, 1.5),
                                                      to_uv=np.diag([3.0, 3.0, 1.0, 1.0])), fov=40.0)
    b.set_camera(64, 40, 60, origin=(0, 2.5, -6), target=(0, 0.5, 0), far=1000)
    b.set_render(integrator, spp=16, max_depth=6, rr_depth=3, seed=4)
    return b


@pytest.mark.parametrize("integrator", ["path", "vol_path"])
def test_textured_projection_matches_oracle(hip_ctx, integrator):
    b = projection_scene(integrator)
    s = hip.Scene.from_builder(hip_ctx, b)
    g, gsq, _ = s.render(b.render_desc(spp=4096, seed=1, flags=F.RAD_RENDER_WANT_SUMSQ), want_sumsq=True)
    of, osq, _ = O.OracleScene(b).render(spp=1024, seed=99)
    z, rel = stats.compare(g, gsq, 4096, of, osq, 1024, tile=8)
    print(f"max |z| {np.abs(z).max():.2f}  relMSE {rel:.2e}")
    assert np.isfinite(g).all() and g.max() > 0
    assert np.abs(z).max() < 4.0
    assert rel < 2e-3


def test_json_textured_projection_end_to_end(hip_ctx, tmp_path):
    import ctypes as C

    b = projection_scene()
    path = b.write_json(str(tmp_path))
    h = C.c_void_p()
    hip.check(hip.lib().rad_scene_load_json(path.encode(), C.byref(h)))
    try:
        d = hip.lib().rad_loaded_scene_desc(h).contents
        rd = hip.lib().rad_loaded_scene_render(h).contents
        fb1 = hip.Scene(hip_ctx, d).render(rd)[0]
        fb2 = hip.Scene.from_builder(hip_ctx, b).render(b.render_desc())[0]
        assert np.array_equal(fb1, fb2)
    finally:
        hip.lib().rad_loaded_scene_free(h)
