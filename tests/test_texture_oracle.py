"""The oracle's textures (texture/bitmap.cpp, texture/chessboard.cpp) and
texture-space footprints (interaction.cpp:83-111) against closed forms:
texel selection of the nearest filter, the reference's bilinear weights
(including its wrap to the last texel below texel 0), the mip chain as a
[1 2 1] tent filter, trilinear level selection, EWA on constant images,
chessboard quadrants, the footprint of a fronto-parallel rectangle, and
constant textures reproducing constant parameters exactly."""
import ctypes
import math

import numpy as np
import pytest

import _oracle as O
from radiantiumplusplus_amd import _ffi as F
from radiantiumplusplus_amd import scenes

IDENT16 = [float(x) for x in np.eye(4, dtype=np.float32).reshape(-1)]


def tex_array(*descs):
    arr = (F.TextureDesc * len(descs))()
    for i, d in enumerate(descs):
        arr[i] = d
    return arr


KEEP = []


def bitmap(img, filt=F.RAD_FILTER_BILINEAR, wrap=F.RAD_WRAP_CLAMP, max_level=0, aniso=0.0):
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


def look(uv, dx=(0, 0), dy=(0, 0)):
    return [uv[0], uv[1], dx[0], dx[1], dy[0], dy[1]]


def ev(arr, idx, rows):
    return O.texture_eval(arr, idx, np.array(rows, np.float32))


def test_nearest_selects_floor_of_u_times_w_minus_1():
    """bitmap.cpp:168-178: texel (floor(u (W - 1)), floor(v (H - 1)))."""
    W, H = 5, 3
    img = (np.arange(H)[:, None] * 10 + np.arange(W)[None, :]).astype(np.float32)
    arr = tex_array(bitmap(img, F.RAD_FILTER_NEAREST))
    rng = np.random.default_rng(1)
    uv = rng.uniform(0, 1, (200, 2))
    out = ev(arr, 0, [look(p) for p in uv])[:, 0]
    want = img[np.floor(uv[:, 1].astype(np.float32) * np.float32(H - 1)).astype(int),
               np.floor(uv[:, 0].astype(np.float32) * np.float32(W - 1)).astype(int)]
    assert np.array_equal(out, want)
    # clamp vs repeat outside [0, 1]
    assert ev(arr, 0, [look((1.7, -0.3))])[0, 0] == img[0, W - 1]
    arr_r = tex_array(bitmap(img, F.RAD_FILTER_NEAREST, F.RAD_WRAP_REPEAT))
    u = np.float32(1.7) - np.floor(np.float32(1.7))
    assert ev(arr_r, 0, [look((1.7, 0.0))])[0, 0] == img[0, int(np.float32(u) * np.float32(W - 1))]


def test_bilinear_weights_and_low_edge_wrap():
    """bitmap.cpp:192-206: at f = u (W - 1) the texel floor(f) is blended with
    the neighbour on the side of the texel centre, weight |f - floor(f) - 0.5|;
    below texel 0 the unsigned neighbour index clamps to the last texel."""
    row = np.array([1.0, 2.0, 4.0, 8.0, 16.0], np.float32)
    img = np.tile(row, (2, 1))
    arr = tex_array(bitmap(img))
    W = len(row)
    f = 2.75  # texel 2, neighbour 3, weight 0.25
    assert ev(arr, 0, [look((f / (W - 1), 0.0))])[0, 0] == pytest.approx(0.75 * 4 + 0.25 * 8, rel=1e-6)
    f = 2.5  # texel centre: exactly the texel
    assert ev(arr, 0, [look((f / (W - 1), 0.0))])[0, 0] == 4.0
    f = 0.25  # below the centre of texel 0: blended with the LAST texel, weight 0.25
    assert ev(arr, 0, [look((f / (W - 1), 0.0))])[0, 0] == pytest.approx(0.75 * 1 + 0.25 * 16, rel=1e-6)
    # u = 1: f = W - 1, the neighbour is W - 2 with weight 0.5
    assert ev(arr, 0, [look((1.0, 0.0))])[0, 0] == pytest.approx(0.5 * 16 + 0.5 * 8, rel=1e-6)


def tent_level(img):
    """One mip level by the [1 2 1] / 4 tent in each axis over texels 2i .. 2i + 2
    (clamped): the mean of four half-way bilinear taps."""
    H, W = img.shape[:2]
    out = np.zeros((max(1, H // 2), max(1, W // 2)) + img.shape[2:], np.float64)
    wts = [0.25, 0.5, 0.25]
    for j in range(out.shape[0]):
        for i in range(out.shape[1]):
            acc = 0
            for a, wy in enumerate(wts):
                for b_, wx in enumerate(wts):
                    acc = acc + wy * wx * img[min(2 * j + a, H - 1), min(2 * i + b_, W - 1)].astype(np.float64)
            out[j, i] = acc
    return out.astype(np.float32)


def test_trilinear_picks_mip_level_of_the_footprint():
    """bitmap.cpp:82-104 with footprint 2^-k: level = L - 1 - k exactly, read
    bilinearly at a texel centre of that level, whose value is the tent-filtered
    mip texel (bitmap.cpp:44-58)."""
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 1, (16, 16)).astype(np.float32)
    levels = [img]
    while levels[-1].shape[0] > 1:
        levels.append(tent_level(levels[-1]))
    L = len(levels)
    assert L == 1 + int(math.log2(16))
    arr = tex_array(bitmap(img, F.RAD_FILTER_TRILINEAR, max_level=-1))
    for k in range(1, L - 1):
        lev = L - 1 - k
        m = levels[lev]
        w = m.shape[1]
        # texel centres (t + 1/2) / (w - 1) lie in [0, 1] for t <= w - 2
        for (ti, tj) in sorted({(0, 0), ((w - 2) // 2, (w - 2) // 3), (w - 2, w - 2)}):
            u = (ti + 0.5) / (w - 1)
            v = (tj + 0.5) / (w - 1)
            got = ev(arr, 0, [look((u, v), (2.0 ** -k, 0), (0, 0))])[0, 0]
            assert got == pytest.approx(float(m[tj, ti]), rel=2e-6, abs=1e-7), (k, ti, tj)
    # below level 0 and beyond the last level: bilinear on level 0 / on the 1x1 level
    assert ev(arr, 0, [look((0.3, 0.6), (1e-4, 0), (0, 0))])[0, 0] == ev(tex_array(bitmap(img)), 0, [look((0.3, 0.6))])[0, 0]
    assert ev(arr, 0, [look((0.3, 0.6), (4.0, 0), (0, 0))])[0, 0] == pytest.approx(float(levels[-1][0, 0]), rel=1e-6)


def test_mip_chain_needs_even_sizes():
    """bitmap.cpp:30-33: odd level-0 sizes disable the chain (trilinear reads level 0)."""
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, (6, 5)).astype(np.float32)
    tri = tex_array(bitmap(img, F.RAD_FILTER_TRILINEAR, max_level=-1))
    bil = tex_array(bitmap(img))
    q = [look((0.37, 0.61), (0.5, 0), (0, 0.5))]
    assert ev(tri, 0, q)[0, 0] == ev(bil, 0, q)[0, 0]


@pytest.mark.parametrize("filt", [F.RAD_FILTER_NEAREST, F.RAD_FILTER_BILINEAR, F.RAD_FILTER_TRILINEAR,
                                  F.RAD_FILTER_ANISOTROPIC])
def test_constant_image_is_constant_under_every_filter(filt):
    img = np.full((8, 16, 3), 0.0, np.float32)
    img[..., :] = [0.25, 0.5, 0.125]  # exact in binary: every blend stays exact
    arr = tex_array(bitmap(img, filt, max_level=-1 if filt >= 2 else 0, aniso=8.0))
    rng = np.random.default_rng(4)
    rows = [look(rng.uniform(-0.5, 1.5, 2), rng.normal(0, 0.05, 2), rng.normal(0, 0.05, 2)) for _ in range(300)]
    out = ev(arr, 0, rows)
    assert np.allclose(out, [0.25, 0.5, 0.125], rtol=2e-6)


def test_ewa_is_a_convex_combination_and_degenerates_to_bilinear():
    """bitmap.cpp:106-134, 214-246: positive weights over level-0 texels; with a
    zero footprint the anisotropic filter is the bilinear one."""
    rng = np.random.default_rng(5)
    img = rng.uniform(0.1, 0.9, (32, 32)).astype(np.float32)
    ani = tex_array(bitmap(img, F.RAD_FILTER_ANISOTROPIC, max_level=-1, aniso=8.0))
    bil = tex_array(bitmap(img))
    rows = [look(rng.uniform(0, 1, 2), rng.normal(0, 0.03, 2), rng.normal(0, 0.003, 2)) for _ in range(300)]
    out = ev(ani, 0, rows)[:, 0]
    assert np.all(out >= img.min() - 1e-6) and np.all(out <= img.max() + 1e-6)
    q = [look((0.41, 0.77)), look((0.9, 0.1))]
    assert np.array_equal(ev(ani, 0, q), ev(bil, 0, q))


def test_chessboard_quadrants_and_remap():
    """chessboard.cpp:33-55: color_a where the fractional parts are on the same
    side of 1/2, color_b elsewhere; remap rescales uv for nested textures."""
    d = F.TextureDesc()
    d.type = F.RAD_TEXTURE_CHESSBOARD
    d.channels = 3
    d.to_uv[:] = IDENT16
    d.color_a[:] = [1, 0, 0]
    d.color_b[:] = [0, 0, 1]
    arr = tex_array(d)
    out = ev(arr, 0, [look((0.25, 0.25)), look((0.75, 0.25)), look((0.25, 0.75)), look((0.75, 0.75)), look((1.25, 0.25))])
    assert np.array_equal(out, [[1, 0, 0], [0, 0, 1], [0, 0, 1], [1, 0, 0], [1, 0, 0]])
    # to_uv scale 4: the pattern repeats every 1/4
    d.to_uv[:] = [float(x) for x in scenes.trs(scale=(4, 4, 1)).reshape(-1)]
    arr = tex_array(d)
    assert np.array_equal(ev(arr, 0, [look((0.0625, 0.0625)), look((0.1875, 0.0625))]), [[1, 0, 0], [0, 0, 1]])
    # nested bitmap with remap: color_b reads the bitmap at 2 * tuv
    ramp = np.linspace(0, 1, 11, dtype=np.float32)[None, :, None].repeat(2, 0).repeat(3, 2)
    nb = bitmap(ramp, F.RAD_FILTER_NEAREST)
    d2 = F.TextureDesc()
    d2.type = F.RAD_TEXTURE_CHESSBOARD
    d2.channels = 3
    d2.to_uv[:] = IDENT16
    d2.color_a[:] = [0.4] * 3
    d2.tex_b = 2
    d2.remap = 1
    arr = tex_array(d2, nb)
    got = ev(arr, 0, [look((0.3, 0.7))])[0]
    want = ev(tex_array(nb), 0, [look((0.6, 1.4))])[0]
    assert np.array_equal(got, want)


def test_footprint_of_a_fronto_parallel_rectangle():
    """interaction.cpp:83-111 for a camera looking straight at a 2s x 2s
    rectangle at distance d (uv = local / 2 + 1/2): |du/dx| = |dv/dy| =
    d tan(fov / 2) / (W s), the cross terms vanish."""
    b = scenes.SceneBuilder()
    s_, dist, W, H, fov = 3.0, 5.0, 80, 60, 50.0
    b.rectangle(to_world=scenes.trs(translate=(0, 0, dist), scale=(s_, s_, 1)), bsdf=b.diffuse(0.5))
    light = b.rectangle(to_world=scenes.trs(translate=(0, 0, -1), scale=(0.1, 0.1, 1)))
    b.area_light(light, 1.0)
    b.set_camera(W, H, fov, origin=(0, 0, 0), target=(0, 0, 1), far=100)
    o = O.OracleScene(b)
    pts = np.array([[40.5, 30.5], [20.25, 10.75], [70.5, 50.5]], np.float32)
    fp = o.camera_footprint(pts)
    want = dist * math.tan(math.radians(fov / 2)) / (W * s_)
    # the pixel step is taken on the near plane, so the footprint is exact only at the centre
    assert abs(abs(fp[0, 2]) - want) < 2e-3 * want
    assert abs(abs(fp[0, 5]) - want) < 2e-3 * want
    assert abs(fp[0, 3]) < 1e-3 * want and abs(fp[0, 4]) < 1e-3 * want
    assert np.all(np.abs(np.abs(fp[:, 2]) - want) < 0.05 * want)


def textured_pair(tex_kind):
    """Two builds of one scene: constant parameters, and the same values as
    1 x 1 nearest-filter textures (which return them unchanged)."""
    out = []
    for textured in (False, True):
        b = scenes.SceneBuilder()
        def c3(v):
            return b.bitmap(np.array(v, np.float32).reshape(1, 1, 3), filter="nearest") if textured else v
        def c1(v):
            return b.bitmap(np.array([[v]], np.float32), filter="nearest") if textured else v
        b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(6, 6, 6)), bsdf=b.diffuse(c3((0.6, 0.5, 0.4))))
        if tex_kind == "metal":
            m = b.rough_metal(alpha=c1(0.3), reflectance=c3((0.9, 0.6, 0.3)))
        elif tex_kind == "plastic":
            m = b.plastic(diffuse=c3((0.2, 0.6, 0.3)), specular=c3((0.9, 0.9, 0.9)))
        else:
            m = b.disney(base_color=c3((0.7, 0.3, 0.2)), roughness=c1(0.4), metallic=c1(0.5))
        b.sphere(center=(0, 1, 0), radius=1, bsdf=m)
        light = b.rectangle(to_world=scenes.trs(translate=(0, 5, -1), axis=(1, 0, 0), angle=90, scale=(1, 1, 1)))
        b.area_light(light, 8.0)
        b.set_camera(24, 16, 55, origin=(0, 3, -8), target=(0, 0.8, 0), far=1000)
        b.set_render("path", spp=4, max_depth=5, rr_depth=3, seed=3)
        out.append(b)
    return out


@pytest.mark.parametrize("kind", ["metal", "plastic", "disney"])
def test_constant_textures_reproduce_constant_parameters(kind):
    """Per-hit parameter lookup (every `_x->Eval(si)` of bsdf/*.cpp) and the
    plastic sampling weight from texel means (plastic.cpp:29-42) leave a scene
    whose textures are constant bit-identical to the untextured one."""
    plain, textured = textured_pair(kind)
    a = O.OracleScene(plain).render(spp=4, seed=7)[0]
    t = O.OracleScene(textured).render(spp=4, seed=7)[0]
    assert np.array_equal(a, t)



def env_furnace(tex_kind):
    """A diffuse sphere inside a uniform-valued environment map."""
    b = scenes.SceneBuilder()
    b.sphere(center=(0, 0, 0), radius=1.0, bsdf=b.diffuse(0.6))
    if tex_kind == "bitmap":
        b.skybox(b.bitmap(np.ones((4, 8, 3), np.float32)))
    else:
        b.skybox(1.0)
    b.set_camera(16, 16, 40, origin=(0, 0, -4), target=(0, 0, 0), far=100)
    b.set_render("path", spp=64, max_depth=-1, rr_depth=1000, seed=7)
    return b


def test_textured_environment_furnace():
    """skybox.cpp with a constant 8 x 4 map and quirks off (continuous in-texel
    sampling): a white furnace, so the sphere shows its albedo and the
    background the map value exactly."""
    b = env_furnace("bitmap")
    fb, _, _ = O.OracleScene(b, quirks=False).render(spp=256, seed=3)
    assert np.allclose(fb[0, 0], 1.0, rtol=1e-6)
    c = fb[6:10, 6:10].mean(axis=(0, 1))
    assert np.allclose(c, 0.6, atol=0.015)


def area_scene(textured, value=4.0):
    b = scenes.SceneBuilder()
    b.rectangle(to_world=scenes.trs(axis=(1, 0, 0), angle=-90, scale=(4, 4, 4)), bsdf=b.diffuse(0.7))
    light = b.rectangle(to_world=scenes.trs(translate=(0, 3, 0), axis=(1, 0, 0), angle=90, scale=(1, 1, 1)))
    rad = b.bitmap(np.full((2, 4, 3), value, np.float32), filter="nearest") if textured else value
    b.area_light(light, rad)
    b.set_camera(24, 16, 60, origin=(0, 2, -5), target=(0, 0, 0), far=100)
    b.set_render("path", spp=64, max_depth=4, rr_depth=100, seed=2)
    return b


def test_textured_area_light_sampling():
    """area.cpp:20-91 with a constant 4 x 2 radiance map: quirks off (texel
    distribution, continuous positions, PdfDirection density) matches the
    constant light statistically; quirks on, the light samples carry density 0
    (SurfaceInteraction::ToDsr) and the floor only receives the MIS-weighted
    BSDF-sampled part, so it is darker."""
    import stats
    c, csq, _ = O.OracleScene(area_scene(False), quirks=False).render(spp=512, seed=1)
    t, tsq, _ = O.OracleScene(area_scene(True), quirks=False).render(spp=512, seed=2)
    z, rel = stats.compare(t, tsq, 512, c, csq, 512, tile=4)
    assert np.abs(z).max() < 4.0, np.abs(z).max()
    q, _, _ = O.OracleScene(area_scene(True), quirks=True).render(spp=512, seed=2)
    floor = slice(10, 16)
    assert q[floor].mean() < 0.9 * c[floor].mean()
