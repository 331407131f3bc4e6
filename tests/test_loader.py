"""Scene JSON loading (BuildContext::SetFromJson / Build mirror) on the host.

The loader reads the reference's on-disk format (scene JSON + OBJ + mitsuba
.vol); the fixtures are written by the scene builder, so the loaded
description must reproduce what the builder holds.  The strongest check: the
oracle renders the loaded scene bit-identically to the builder's scene.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import _oracle as O
from radiantiumplusplus_amd import _ffi as F
from radiantiumplusplus_amd import hip, scenes


def load(path):
    h = C.c_void_p()
    st = hip.lib().rad_scene_load_json(path.encode(), C.byref(h))
    if st != 0:
        raise hip.RadError(st, hip.lib().rad_hip_last_error().decode())
    return h


def desc_of(h):
    return hip.lib().rad_loaded_scene_desc(h).contents


def test_many_balls_round_trip(tmp_path):
    b = scenes.many_balls(width=32, height=24, spp=4, n_spheres=5, slices=12)
    path = b.write_json(str(tmp_path))
    h = load(path)
    try:
        d = desc_of(h)
        ref = b.to_desc()
        assert d.shape_count == ref.shape_count
        assert d.light_count == ref.light_count == 1
        assert d.camera.width == 32 and d.camera.height == 24
        assert np.allclose(list(d.camera.to_world), list(ref.camera.to_world))
        for i in range(d.shape_count):
            s, r = d.shapes[i], ref.shapes[i]
            assert s.type == r.type
            assert list(s.to_world) == list(r.to_world)
            assert s.triangle_count == r.triangle_count
            if s.type == F.RAD_SHAPE_MESH:
                # OBJ loading deduplicates vertices, so compare triangle corners
                n = s.triangle_count
                pl = np.ctypeslib.as_array(s.positions, (s.vertex_count * 3,)).reshape(-1, 3)
                il = np.ctypeslib.as_array(s.indices, (n * 3,))
                pr = np.ctypeslib.as_array(r.positions, (r.vertex_count * 3,)).reshape(-1, 3)
                ir = np.ctypeslib.as_array(r.indices, (n * 3,))
                assert np.array_equal(pl[il], pr[ir])
            assert (s.light >= 0) == (r.light >= 0)
        for i in range(d.bsdf_count):
            assert d.bsdfs[i].type == F.RAD_BSDF_DIFFUSE
        r = hip.lib().rad_loaded_scene_render(h).contents
        assert r.spp == 4 and r.max_depth == 8 and r.rr_depth == 3
        assert hip.lib().rad_loaded_scene_renderer_type(h) == b"path"
        assert hip.lib().rad_loaded_scene_device(h) == b"hip"
    finally:
        hip.lib().rad_loaded_scene_free(h)


def test_loaded_scene_renders_identically_in_oracle(tmp_path):
    b = scenes.many_balls(width=24, height=24, spp=4, n_spheres=6, slices=12)
    path = b.write_json(str(tmp_path))
    h = load(path)
    try:
        d = desc_of(h)
        s1 = O.OracleScene(b)
        fb1, _, _ = s1.render(threads=2)

        class Wrapped:
            """Builder-shaped view of the loaded description for the oracle."""
            camera = b.camera
            renderer = b.renderer
            sampler = b.sampler

            def to_desc(self):
                return d

        s2 = O.OracleScene(Wrapped())
        fb2, _, _ = s2.render(threads=2)
        assert np.array_equal(fb1, fb2)
    finally:
        hip.lib().rad_loaded_scene_free(h)


def _write_scene(tmp_path, scene):
    p = tmp_path / "s.json"
    p.write_text(json.dumps(scene))
    return str(p)


BASE = {
    "camera": {"type": "perspective", "fov": 40, "origin": [0, 1, -5], "target": [0, 0, 0], "up": [0, 1, 0],
               "resolution": [16, 8]},
    "renderer": {"type": "path"},
}


def test_defaults_and_lookat(tmp_path):
    sc = dict(BASE)
    sc["scene"] = [{"shape": {"type": "rectangle"}, "light": {"type": "area", "radiance": 3}}]
    h = load(_write_scene(tmp_path, sc))
    try:
        d = desc_of(h)
        m = np.array(list(d.camera.to_world)).reshape(4, 4)
        assert np.allclose(m, scenes.look_at_left_hand((0, 1, -5), (0, 0, 0), (0, 1, 0)), atol=1e-6)
        assert d.camera.near_clip == pytest.approx(0.001) and d.camera.far_clip == 100
        # a light entity without a BSDF receives a default diffuse (build_context.cpp:233-236)
        assert d.shapes[0].bsdf == 0 and d.bsdfs[0].type == F.RAD_BSDF_DIFFUSE
        assert list(d.bsdfs[0].reflectance) == pytest.approx([0.5, 0.5, 0.5])
        assert list(d.lights[0].radiance) == [3, 3, 3]
        r = hip.lib().rad_loaded_scene_render(h).contents
        assert r.spp == 16 and r.max_depth == -1 and r.rr_depth == 3
    finally:
        hip.lib().rad_loaded_scene_free(h)


def test_children_compose_child_times_parent(tmp_path):
    """Child to_world = Child * Parent (build_context.cpp:160)."""
    sc = dict(BASE)
    parent = {"to_world": {"translate": [1, 0, 0]},
              "children": [{"to_world": {"rotate": {"axis": [0, 0, 1], "angle": 90}},
                            "shape": {"type": "sphere"}, "bsdf": {"type": "diffuse"}}],
              "shape": {"type": "rectangle"}, "light": {"type": "area"}}
    sc["scene"] = [parent]
    h = load(_write_scene(tmp_path, sc))
    try:
        d = desc_of(h)
        assert d.shape_count == 2
        # BFS: the parent's shape first, then the child
        assert d.shapes[0].type == F.RAD_SHAPE_RECTANGLE and d.shapes[1].type == F.RAD_SHAPE_SPHERE
        child = np.array(list(d.shapes[1].to_world)).reshape(4, 4)
        expect = scenes.trs(axis=(0, 0, 1), angle=90) @ scenes.trs(translate=(1, 0, 0))
        assert np.allclose(child, expect, atol=1e-6)
        # parent applied first: the translation is rotated by the child
        assert np.allclose(child @ np.array([0, 0, 0, 1]), [0, 1, 0, 1], atol=1e-6)
    finally:
        hip.lib().rad_loaded_scene_free(h)


def test_bsdf_vars_and_disney_flags(tmp_path):
    sc = dict(BASE)
    sc["bsdf_vars"] = [{"name": "gold", "bsdf": {"type": "rough_metal", "alpha": 0.2}}]
    sc["scene"] = [
        {"shape": {"type": "sphere"}, "bsdf": {"ref": "gold"}},
        {"shape": {"type": "cube"}, "bsdf": {"type": "disney", "metallic": 0.0, "sheen": 0.5, "specular": 0.3}},
        {"shape": {"type": "rectangle"}, "light": {"type": "area"}},
    ]
    h = load(_write_scene(tmp_path, sc))
    try:
        d = desc_of(h)
        assert d.bsdfs[d.shapes[0].bsdf].type == F.RAD_BSDF_ROUGH_METAL
        assert d.bsdfs[d.shapes[0].bsdf].alpha_u == pytest.approx(0.2)
        dis = d.bsdfs[d.shapes[1].bsdf]
        # metallic given as 0 does not enable the lobe (disney.cpp:459-471)
        assert not (dis.disney_flags & F.DISNEY_FLAGS["metallic"])
        assert dis.disney_flags & F.DISNEY_FLAGS["sheen"]
        assert dis.specular == pytest.approx(0.3)
    finally:
        hip.lib().rad_loaded_scene_free(h)


@pytest.mark.parametrize("bad,code", [
    ({"type": "plastic", "diffuse": {"type": "chessboard"}}, 4),  # textures: RAD_ERR_UNSUPPORTED
    ({"type": "two_side", "out_side": {"type": "diffuse"}}, 4),   # the reference's broken form
    ({"type": "mask"}, 1),                                         # no nested bsdf
    ({"type": "nonexistent"}, 1),                                  # unknown factory -> RAD_ERR_INVALID_ARG
])
def test_unsupported_and_unknown_bsdfs_are_rejected(tmp_path, bad, code):
    sc = dict(BASE)
    sc["scene"] = [{"shape": {"type": "sphere"}, "bsdf": bad}, {"shape": {"type": "rectangle"}, "light": {"type": "area"}}]
    with pytest.raises(hip.RadError) as e:
        load(_write_scene(tmp_path, sc))
    assert e.value.status == code


def test_volume_assets(tmp_path):
    b = scenes.volume_box(width=8, height=8, spp=1, res=8)
    path = b.write_json(str(tmp_path))
    h = load(path)
    try:
        d = desc_of(h)
        assert d.medium_count == 1
        m = d.media[0]
        assert m.type == F.RAD_MEDIUM_HETEROGENEOUS and list(m.grid_size) == [8, 8, 8]
        g = np.ctypeslib.as_array(m.grid, (512,))
        assert np.array_equal(g, b.media[0]["grid"].reshape(-1))
        assert m.phase == F.RAD_PHASE_HG and m.g == pytest.approx(0.3)
        cube = [s for s in d.shapes[: d.shape_count] if s.type == F.RAD_SHAPE_CUBE][0]
        assert cube.bsdf == -1 and cube.medium_inside == 0
        assert hip.lib().rad_loaded_scene_renderer_type(h) == b"vol_path"
        assert hip.lib().rad_loaded_scene_render(h).contents.rr_depth == 5
    finally:
        hip.lib().rad_loaded_scene_free(h)


def test_exr_writer(tmp_path):
    img = np.random.default_rng(0).random((5, 7, 3), dtype=np.float32)
    p = str(tmp_path / "o.exr")
    assert hip.lib().rad_write_exr(p.encode(), img.ctypes.data, 7, 5) == 0
    raw = open(p, "rb").read()
    assert raw[:4] == (20000630).to_bytes(4, "little")
    # last scanline block holds B, G, R planes of row 4
    tail = np.frombuffer(raw[-7 * 3 * 4:], np.float32).reshape(3, 7)
    assert np.array_equal(tail[2], img[4, :, 0]) and np.array_equal(tail[0], img[4, :, 2])


def test_remaining_bsdfs_and_lights_load_with_reference_defaults(tmp_path):
    """plastic.cpp / rough_plastic.cpp / perfect_*.cpp / two_side.cpp /
    mask.cpp / point.cpp / projection.cpp defaults, nested BSDFs appended
    and referenced by index."""
    sc = dict(BASE)
    sc["scene"] = [
        {"shape": {"type": "sphere"}, "bsdf": {"type": "plastic"}},
        {"shape": {"type": "sphere"}, "bsdf": {"type": "rough_plastic", "alpha": 0.3, "distribution": "beckmann"}},
        {"shape": {"type": "sphere"}, "bsdf": {"type": "perfect_mirror"}},
        {"shape": {"type": "sphere"}, "bsdf": {"type": "perfect_glass"}},
        {"shape": {"type": "sphere"}, "bsdf": {"type": "mask", "opacity": 0.25,
                                               "nested": {"type": "two_side", "two_side": {"type": "diffuse"}}}},
        {"to_world": {"translate": [1, 2, 3]}, "light": {"type": "point", "position": [0, 1, 0], "intensity": 4}},
        {"light": {"type": "projection", "irradiance": [1, 2, 3], "scale": 2, "fov": 30}},
    ]
    h = load(_write_scene(tmp_path, sc))
    try:
        d = hip.lib().rad_loaded_scene_desc(h).contents
        b = [d.bsdfs[i] for i in range(d.bsdf_count)]
        types = [x.type for x in b]
        assert types[:4] == [F.RAD_BSDF_PLASTIC, F.RAD_BSDF_ROUGH_PLASTIC, F.RAD_BSDF_PERFECT_MIRROR,
                             F.RAD_BSDF_PERFECT_GLASS]
        p = b[0]
        assert p.int_ior == pytest.approx(1.49) and p.ext_ior == pytest.approx(1.000277)
        assert list(p.diffuse) == pytest.approx([0.5] * 3) and list(p.specular_reflectance) == pytest.approx([1] * 3)
        assert b[1].alpha_u == pytest.approx(0.3) and b[1].distribution == F.RAD_MICROFACET_BECKMANN
        assert list(b[2].eta) == [0, 0, 0] and list(b[2].k) == [1, 1, 1]
        assert b[3].int_ior == pytest.approx(1.5046)
        shape_bsdf = d.shapes[4].bsdf
        m = b[shape_bsdf]
        assert m.type == F.RAD_BSDF_MASK and m.opacity == pytest.approx(0.25)
        ts = b[m.nested]
        assert ts.type == F.RAD_BSDF_TWO_SIDE and b[ts.nested].type == F.RAD_BSDF_DIFFUSE
        lights = [d.lights[i] for i in range(d.light_count)]
        pt = [l for l in lights if l.type == F.RAD_LIGHT_POINT][0]
        assert [pt.to_world[3], pt.to_world[7], pt.to_world[11]] == pytest.approx([1, 3, 3])
        assert list(pt.radiance) == pytest.approx([4] * 3)
        pj = [l for l in lights if l.type == F.RAD_LIGHT_PROJECTION][0]
        assert list(pj.radiance) == pytest.approx([2, 4, 6]) and pj.fov == pytest.approx(30)
    finally:
        hip.lib().rad_loaded_scene_free(h)
