"""Seeded scene generators and the scene builder.

A `SceneBuilder` records shapes, BSDFs, lights, media and the camera exactly
as the reference's BuildContext would create them from a scene JSON
(module/rad.offline/src/build/build_context.cpp:88-321), and can emit
  * the plain-C `rad_scene_desc` the library consumes (`to_desc`), and
  * the reference's on-disk format: scene JSON + OBJ meshes + mitsuba `.vol`
    grids (`write_json`), which the library's JSON loader reads back.

Generators for the benchmark scenes ("many balls", the 8.4M-triangle terrain,
the heterogeneous-medium box, the door-gap guiding room) are deterministic in
their seed.  Procedural meshes follow TriangleModel::CreateSphere and
CreateGrid (module/rad.core/src/triangle_model.cpp:53-113, 478-521).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from . import _ffi as F

IDENTITY = np.eye(4, dtype=np.float32)


# --------------------------------------------------------------------------- transforms
def rotation_matrix(axis, angle_deg):
    """Eigen::AngleAxis(Radian(angle), axis).toRotationMatrix() (config_node_ext.cpp:7-13)."""
    x, y, z = (float(a) for a in axis)
    a = math.radians(float(angle_deg))
    s, c = math.sin(a), math.cos(a)
    t = 1.0 - c
    return np.array(
        [
            [t * x * x + c, t * x * y - s * z, t * x * z + s * y],
            [t * x * y + s * z, t * y * y + c, t * y * z - s * x],
            [t * x * z - s * y, t * y * z + s * x, t * z * z + c],
        ],
        dtype=np.float64,
    )


def trs(translate=(0, 0, 0), axis=(0, 1, 0), angle=0.0, scale=(1, 1, 1)):
    """Translation * rotation * scale, the object form of `to_world` (config_node_ext.cpp:52-71)."""
    m = np.eye(4, dtype=np.float64)
    m[:3, :3] = rotation_matrix(axis, angle) @ np.diag(np.asarray(scale, dtype=np.float64))
    m[:3, 3] = np.asarray(translate, dtype=np.float64)
    return m.astype(np.float32)


def look_at_left_hand(origin, target, up):
    """math_ext.h:65-75 (float arithmetic)."""
    o = np.asarray(origin, np.float32)
    t = np.asarray(target, np.float32)
    u = np.asarray(up, np.float32)
    d = t - o
    d = d / np.float32(np.linalg.norm(d))
    left = np.cross(u / np.float32(np.linalg.norm(u)), d).astype(np.float32)
    left = left / np.float32(np.linalg.norm(left))
    real_up = np.cross(d, left).astype(np.float32)
    real_up = real_up / np.float32(np.linalg.norm(real_up))
    m = np.eye(4, dtype=np.float32)
    m[:3, 0] = left
    m[:3, 1] = real_up
    m[:3, 2] = d
    m[:3, 3] = o
    return m


# --------------------------------------------------------------------------- procedural meshes
def create_sphere(radius, slices):
    """TriangleModel::CreateSphere (triangle_model.cpp:53-113), float32 arithmetic."""
    slices = int(slices)
    parallels = slices // 2
    step = np.float32(2.0 * np.pi) / np.float32(slices)
    i = np.arange(parallels + 1, dtype=np.float32)[:, None]
    j = np.arange(slices + 1, dtype=np.float32)[None, :]
    r = np.float32(radius)
    si, ci = np.sin(step * i), np.cos(step * i)
    sj, cj = np.sin(step * j), np.cos(step * j)
    px = r * si * sj
    py = np.broadcast_to(r * ci, px.shape)
    pz = r * si * cj
    pos = np.stack([px, py, pz], -1).reshape(-1, 3).astype(np.float32)
    nrm = (pos / r).astype(np.float32)
    tx = np.broadcast_to(j / np.float32(slices), px.shape)
    ty = np.broadcast_to(np.float32(1) - i / np.float32(parallels), px.shape)
    uv = np.stack([tx, ty], -1).reshape(-1, 2).astype(np.float32)
    ii = np.arange(parallels, dtype=np.uint32)[:, None]
    jj = np.arange(slices, dtype=np.uint32)[None, :]
    w = np.uint32(slices + 1)
    a = ii * w + jj
    b = (ii + 1) * w + jj
    c = (ii + 1) * w + (jj + 1)
    d = ii * w + (jj + 1)
    idx = np.stack([a, b, c, a, c, d], -1).reshape(-1).astype(np.uint32)
    return pos, nrm, uv, idx


def create_grid(width, depth, m, n):
    """TriangleModel::CreateGrid (triangle_model.cpp:478-521): m rows (z), n columns (x)."""
    hw, hd = np.float32(0.5 * width), np.float32(0.5 * depth)
    dx = np.float32(width) / np.float32(n - 1)
    dz = np.float32(depth) / np.float32(m - 1)
    i = np.arange(m, dtype=np.float32)[:, None]
    j = np.arange(n, dtype=np.float32)[None, :]
    z = np.broadcast_to(hd - i * dz, (m, n))
    x = np.broadcast_to(-hw + j * dx, (m, n))
    pos = np.stack([x, np.zeros((m, n), np.float32), z], -1).reshape(-1, 3).astype(np.float32)
    nrm = np.tile(np.array([0, 1, 0], np.float32), (m * n, 1))
    uv = np.stack(
        [np.broadcast_to(j / np.float32(n - 1), (m, n)), np.broadcast_to(i / np.float32(m - 1), (m, n))], -1
    ).reshape(-1, 2).astype(np.float32)
    ii = np.arange(m - 1, dtype=np.uint32)[:, None]
    jj = np.arange(n - 1, dtype=np.uint32)[None, :]
    nn = np.uint32(n)
    v00 = ii * nn + jj
    v01 = ii * nn + jj + 1
    v10 = (ii + 1) * nn + jj
    v11 = (ii + 1) * nn + jj + 1
    idx = np.stack([v00, v01, v10, v10, v01, v11], -1).reshape(-1).astype(np.uint32)
    return pos, nrm, uv, idx


def _value_noise_grid(rng, res):
    return rng.random((res + 1, res + 1), dtype=np.float64) * 2.0 - 1.0


def fbm_heightfield(m, n, octaves, amplitude, seed):
    """Seeded fractal Brownian motion of smooth value noise on an m x n lattice."""
    rng = np.random.default_rng(seed)
    u = np.linspace(0.0, 1.0, n, dtype=np.float64)[None, :]
    v = np.linspace(0.0, 1.0, m, dtype=np.float64)[:, None]
    h = np.zeros((m, n), np.float64)
    amp, freq, norm = 1.0, 4, 0.0
    for _ in range(octaves):
        g = _value_noise_grid(rng, freq)
        fu, fv = u * freq, v * freq
        iu = np.minimum(fu.astype(np.int64), freq - 1)
        iv = np.minimum(fv.astype(np.int64), freq - 1)
        tu, tv = fu - iu, fv - iv
        su, sv = tu * tu * (3 - 2 * tu), tv * tv * (3 - 2 * tv)
        a = g[iv, iu]
        b = g[iv, iu + 1]
        c = g[iv + 1, iu]
        d = g[iv + 1, iu + 1]
        h += amp * ((a * (1 - su) + b * su) * (1 - sv) + (c * (1 - su) + d * su) * sv)
        norm += amp
        amp *= 0.5
        freq *= 2
    return (h / norm * amplitude).astype(np.float32)


def perlin3d(res, octaves, seed):
    """Seeded gradient (Perlin) noise fBm on a res^3 lattice, x fastest in memory.

    Per octave the noise is a sum over the 8 lattice corners of quintic-fade
    weights times gradient . offset; with the per-axis weights written as
    sparse res x (freq + 1) matrices (W: fade weight, P: fade weight times the
    offset along that axis) each gradient component contracts separably,
    n[z, y, x] = sum_c sum_kji Mz[z, k] My[y, j] Mx[x, i] g[k, j, i, c]."""
    rng = np.random.default_rng(seed)
    coords = (np.arange(res, dtype=np.float64) + 0.5) / res
    out = np.zeros((res, res, res), np.float64)  # [z, y, x]
    amp, freq = 1.0, 4
    ar = np.arange(res)

    def contract(G, Mz, My, Mx):
        T = np.tensordot(G, Mx, axes=([2], [1]))  # (k, j, x)
        T = np.tensordot(T, My, axes=([1], [1]))  # (k, x, y)
        return np.tensordot(Mz, T, axes=([1], [0])).transpose(0, 2, 1)  # (z, y, x)

    for _ in range(octaves):
        g = rng.normal(size=(freq + 1, freq + 1, freq + 1, 3))
        g /= np.linalg.norm(g, axis=-1, keepdims=True)
        f = coords * freq
        i = np.minimum(f.astype(np.int64), freq - 1)
        t = f - i
        s = t * t * t * (t * (t * 6 - 15) + 10)
        W = np.zeros((res, freq + 1))
        W[ar, i], W[ar, i + 1] = 1 - s, s
        P = np.zeros((res, freq + 1))
        P[ar, i], P[ar, i + 1] = (1 - s) * t, s * (t - 1)
        # g[..., 0] pairs with the x offset, 1 with y, 2 with z
        out += amp * (contract(g[..., 0], W, W, P) + contract(g[..., 1], W, P, W) + contract(g[..., 2], P, W, W))
        amp *= 0.5
        freq *= 2
    return out.astype(np.float32)


# --------------------------------------------------------------------------- builder
def _f3(v):
    if isinstance(v, dict):  # a texture (SceneBuilder.bitmap / chessboard)
        return v
    if np.isscalar(v):
        return [float(v)] * 3
    return [float(a) for a in v]


def _f1(v):
    return v if isinstance(v, dict) else float(v)


_FILTERS = {"nearest": 0, "bilinear": 1, "trilinear": 2, "anisotropic": 3}
# BSDF parameter -> (rad_bsdf_desc.tex slot, colour?)
_TEX_SLOTS = {
    "reflectance": (0, True), "transmittance": (1, True), "eta": (2, True), "k": (3, True),
    "alpha_u": (4, False), "alpha_v": (5, False), "base_color": (6, True), "roughness": (7, False),
    "anisotropic": (8, False), "spec_trans": (9, False), "sheen": (10, False), "sheen_tint": (11, False),
    "flatness": (12, False), "spec_tint": (13, False), "metallic": (14, False), "clearcoat": (15, False),
    "clearcoat_gloss": (16, False), "diffuse": (17, True), "specular": (18, True), "opacity": (19, False),
}


@dataclass
class _Shape:
    kind: str
    to_world: np.ndarray
    bsdf: int = -1
    light: int = -1
    medium_inside: int = -1
    medium_outside: int = -1
    flip_normal: bool = False
    mesh: tuple | None = None  # positions, normals, uvs, indices
    center: tuple = (0.0, 0.0, 0.0)
    radius: float = 1.0
    asset: str | None = None
    json_shape: dict = field(default_factory=dict)


class SceneBuilder:
    def __init__(self):
        self.shapes: list[_Shape] = []
        self.bsdfs: list[dict] = []
        self.lights: list[dict] = []
        self.media: list[dict] = []
        self.global_medium = -1
        self.camera = None
        self.renderer = {"type": "path", "max_depth": -1, "rr_depth": 3}
        self.sampler = {"type": "independent", "sample_count": 16, "seed": 0}
        self._keep = []

    # -- materials -------------------------------------------------------
    def diffuse(self, reflectance=0.5):
        self.bsdfs.append({"type": "diffuse", "reflectance": _f3(reflectance)})
        return len(self.bsdfs) - 1

    def rough_metal(self, alpha=0.1, reflectance=1.0, eta=(0.1999899, 0.9220823, 1.0998751),
                    k=(3.9046354, 2.4476316, 2.1376500), distribution="ggx", sample_visible=True):
        self.bsdfs.append({"type": "rough_metal", "alpha": _f1(alpha), "reflectance": _f3(reflectance),
                           "eta": _f3(eta), "k": _f3(k), "distribution": distribution,
                           "sample_visible": bool(sample_visible)})
        return len(self.bsdfs) - 1

    def rough_glass(self, alpha=0.1, int_ior=1.5046, ext_ior=1.000277, reflectance=1.0, transmittance=1.0,
                    distribution="ggx", sample_visible=True):
        self.bsdfs.append({"type": "rough_glass", "alpha": _f1(alpha), "int_ior": float(int_ior),
                           "ext_ior": float(ext_ior), "reflectance": _f3(reflectance),
                           "transmittance": _f3(transmittance), "distribution": distribution,
                           "sample_visible": bool(sample_visible)})
        return len(self.bsdfs) - 1

    def disney(self, **params):
        d = {"type": "disney"}
        for k, v in params.items():
            d[k] = _f3(v) if k == "base_color" else _f1(v)
        self.bsdfs.append(d)
        return len(self.bsdfs) - 1

    def plastic(self, diffuse=0.5, specular=1.0, int_ior=1.49, ext_ior=1.000277, nonlinear=False):
        self.bsdfs.append({"type": "plastic", "diffuse": _f3(diffuse), "specular": _f3(specular),
                           "int_ior": float(int_ior), "ext_ior": float(ext_ior), "no_linear": bool(nonlinear)})
        return len(self.bsdfs) - 1

    def rough_plastic(self, alpha=0.1, diffuse=0.5, specular=1.0, int_ior=1.49, ext_ior=1.000277, nonlinear=False,
                      distribution="ggx", sample_visible=True):
        self.bsdfs.append({"type": "rough_plastic", "alpha": _f1(alpha), "diffuse": _f3(diffuse),
                           "specular": _f3(specular), "int_ior": float(int_ior), "ext_ior": float(ext_ior),
                           "no_linear": bool(nonlinear), "distribution": distribution,
                           "sample_visible": bool(sample_visible)})
        return len(self.bsdfs) - 1

    def perfect_mirror(self, reflectance=1.0, eta=0.0, k=1.0):
        self.bsdfs.append({"type": "perfect_mirror", "reflectance": _f3(reflectance), "eta": _f3(eta), "k": _f3(k)})
        return len(self.bsdfs) - 1

    def perfect_glass(self, int_ior=1.5046, ext_ior=1.000277, reflectance=1.0, transmittance=1.0):
        self.bsdfs.append({"type": "perfect_glass", "int_ior": float(int_ior), "ext_ior": float(ext_ior),
                           "reflectance": _f3(reflectance), "transmittance": _f3(transmittance)})
        return len(self.bsdfs) - 1

    def two_side(self, inner):
        """`two_side` with the same BSDF on both faces (the "two_side" key form)."""
        self.bsdfs.append({"type": "two_side", "two_side": dict(self.bsdfs[inner])})
        return len(self.bsdfs) - 1

    def mask(self, inner, opacity=0.5):
        self.bsdfs.append({"type": "mask", "opacity": _f1(opacity), "nested": dict(self.bsdfs[inner])})
        return len(self.bsdfs) - 1

    # -- textures --------------------------------------------------------
    @staticmethod
    def bitmap(texels, filter="bilinear", wrap="clamp", max_level=None, anisotropic_level=None, name=None):
        """A `bitmap` texture over linear texels (H x W x 3 colour, or H x W mono),
        texel row 0 being the image's bottom row (the reference's flipped load)."""
        t = np.ascontiguousarray(texels, np.float32)
        if t.ndim == 3 and t.shape[2] == 1:
            t = t[:, :, 0]
        return {"type": "bitmap", "_texels": t, "filter": filter, "wrap": wrap, "max_level": max_level,
                "anisotropic_level": anisotropic_level, "asset_name": name}

    @staticmethod
    def chessboard(color_a=0.4, color_b=0.2, to_uv=None, remap=False):
        """A `chessboard` texture; color_a / color_b are constants or bitmaps."""
        return {"type": "chessboard", "color_a": color_a, "color_b": color_b,
                "to_uv": None if to_uv is None else np.asarray(to_uv, np.float32), "remap": bool(remap)}

    def _texture(self, tex, colour, out):
        """Append the TextureDesc of `tex` (and of the bitmaps it nests) to out; 1-based index."""
        td = F.TextureDesc()
        td.channels = 3 if colour else 1
        td.to_uv[:] = [float(x) for x in IDENTITY.reshape(-1)]
        if tex["type"] == "bitmap":
            t = tex["_texels"]
            if colour != (t.ndim == 3):
                raise ValueError("bitmap channel count does not match the parameter")
            if colour and t.shape[2] != 3:
                raise ValueError("colour bitmaps need 3 channels")
            td.type = F.RAD_TEXTURE_BITMAP
            td.height, td.width = t.shape[0], t.shape[1]
            td.texels = self._ptr(t, C.c_float)
            td.filter = _FILTERS[tex["filter"]]
            td.wrap = F.RAD_WRAP_REPEAT if tex["wrap"] == "repeat" else F.RAD_WRAP_CLAMP
            mip = tex["filter"] in ("trilinear", "anisotropic")
            td.max_level = tex["max_level"] if tex["max_level"] is not None else (-1 if mip else 0)
            aniso = tex["anisotropic_level"]
            td.anisotropic_level = aniso if aniso is not None else (8.0 if tex["filter"] == "anisotropic" else 0.0)
        else:
            td.type = F.RAD_TEXTURE_CHESSBOARD
            if tex["to_uv"] is not None:
                td.to_uv[:] = [float(x) for x in tex["to_uv"].reshape(-1)]
            td.remap = 1 if tex["remap"] else 0
            for key, slot, cfield in (("color_a", "tex_a", "color_a"), ("color_b", "tex_b", "color_b")):
                v = tex[key]
                if isinstance(v, dict):
                    setattr(td, slot, self._texture(v, colour, out))
                else:
                    getattr(td, cfield)[:] = _f3(v) if colour else [float(v), 0.0, 0.0]
        out.append(td)
        return len(out)

    # -- media -----------------------------------------------------------
    def homogeneous(self, sigma_t=1.0, albedo=0.75, scale=1.0, g=None):
        phase = {"type": "isotropic_phase"} if g is None else {"type": "hg_phase", "g": float(g)}
        self.media.append({"type": "homogeneous", "sigma_t": _f3(sigma_t), "albedo": _f3(albedo),
                           "scale": float(scale), "phase": phase, "to_world": IDENTITY})
        return len(self.media) - 1

    def heterogeneous(self, grid, albedo=0.75, scale=1.0, g=None, to_world=IDENTITY, name="density"):
        grid = np.ascontiguousarray(grid, np.float32)
        phase = {"type": "isotropic_phase"} if g is None else {"type": "hg_phase", "g": float(g)}
        self.media.append({"type": "heterogeneous", "grid": grid, "grid_name": name, "albedo": _f3(albedo),
                           "scale": float(scale), "phase": phase, "to_world": np.asarray(to_world, np.float32)})
        return len(self.media) - 1

    # -- shapes ----------------------------------------------------------
    def mesh(self, positions, indices, normals=None, uvs=None, to_world=IDENTITY, bsdf=-1, name=None, **kw):
        positions = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        indices = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        normals = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        uvs = None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
        name = name or f"mesh{len(self.shapes)}"
        s = _Shape("mesh", np.asarray(to_world, np.float32), bsdf=bsdf, mesh=(positions, normals, uvs, indices),
                   asset=name, **kw)
        self.shapes.append(s)
        return len(self.shapes) - 1

    def sphere(self, center=(0, 0, 0), radius=1.0, to_world=IDENTITY, bsdf=-1, **kw):
        s = _Shape("sphere", np.asarray(to_world, np.float32), bsdf=bsdf, center=tuple(float(c) for c in center),
                   radius=float(radius), **kw)
        self.shapes.append(s)
        return len(self.shapes) - 1

    def rectangle(self, to_world=IDENTITY, bsdf=-1, flip_normal=False, **kw):
        s = _Shape("rectangle", np.asarray(to_world, np.float32), bsdf=bsdf, flip_normal=flip_normal, **kw)
        self.shapes.append(s)
        return len(self.shapes) - 1

    def cube(self, to_world=IDENTITY, bsdf=-1, **kw):
        s = _Shape("cube", np.asarray(to_world, np.float32), bsdf=bsdf, **kw)
        self.shapes.append(s)
        return len(self.shapes) - 1

    def area_light(self, shape, radiance):
        """Attach an `area` light; a light entity without a BSDF gets a default
        diffuse (build_context.cpp:233-236)."""
        self.lights.append({"type": "area", "radiance": _f3(radiance), "shape": shape})
        li = len(self.lights) - 1
        self.shapes[shape].light = li
        if self.shapes[shape].bsdf < 0:
            self.shapes[shape].bsdf = self.diffuse(0.5)
            self.shapes[shape].json_shape["auto_bsdf"] = True
        return li

    def point_light(self, position, intensity=1.0):
        """`point` light (light/point.cpp) at a world position."""
        tw = np.eye(4, dtype=np.float32)
        tw[:3, 3] = position
        self.lights.append({"type": "point", "radiance": _f3(intensity), "shape": -1, "to_world": tw,
                            "position": [float(x) for x in position]})
        return len(self.lights) - 1

    def projection_light(self, to_world, irradiance=1.0, scale=1.0, fov=90.0):
        """`projection` light (light/projection.cpp) placed by to_world and
        projecting along its +z.  `irradiance` is a constant or an RGB texture
        (`bitmap` / `chessboard` dict), whose width / height sets the aspect."""
        sc = _f3(scale)
        if isinstance(irradiance, dict):
            irr, rad = irradiance, irradiance
        else:
            irr = _f3(irradiance)
            rad = [float(np.float32(a) * np.float32(b)) for a, b in zip(irr, sc)]
        self.lights.append({"type": "projection", "radiance": rad, "irradiance": irr, "scale": sc,
                            "fov": float(fov), "shape": -1, "to_world": np.asarray(to_world, np.float32)})
        return len(self.lights) - 1

    def skybox(self, radiance, to_world=IDENTITY):
        self.lights.append({"type": "skybox", "radiance": _f3(radiance), "shape": -1,
                            "to_world": np.asarray(to_world, np.float32)})
        return len(self.lights) - 1

    def set_camera(self, width, height, fov, origin=None, target=None, up=(0, 1, 0), to_world=None,
                   near=0.001, far=100.0):
        if to_world is None:
            to_world = look_at_left_hand(origin, target, up)
        self.camera = dict(width=int(width), height=int(height), fov=float(fov), near=float(near), far=float(far),
                           to_world=np.asarray(to_world, np.float32),
                           json=dict(origin=origin, target=target, up=up) if origin is not None else None)

    def set_render(self, integrator="path", spp=16, max_depth=-1, rr_depth=None, seed=0):
        if rr_depth is None:
            rr_depth = 3 if integrator == "path" else 5
        self.renderer = {"type": integrator, "max_depth": int(max_depth), "rr_depth": int(rr_depth)}
        self.sampler = {"type": "independent", "sample_count": int(spp), "seed": int(seed)}

    @property
    def triangle_count(self):
        return sum(int(s.mesh[3].size // 3) for s in self.shapes if s.kind == "mesh") + \
            12 * sum(1 for s in self.shapes if s.kind == "cube")

    # -- plain-C description -------------------------------------------------
    def _ptr(self, arr, typ):
        if arr is None:
            return C.cast(None, C.POINTER(typ))
        arr = np.ascontiguousarray(arr)
        self._keep.append(arr)
        return arr.ctypes.data_as(C.POINTER(typ))

    def bsdf_desc(self, d, textures=None):
        """BsdfDesc of one BSDF dict; textured parameters are appended to
        `textures` (a list of TextureDesc) and referenced by index."""
        textured = {}
        consts = {}
        for key, v in d.items():
            if isinstance(v, dict) and v.get("type") in ("bitmap", "chessboard"):
                textured[key] = v
            else:
                consts[key] = v
        b = self._bsdf_consts(consts)
        for key, tex in textured.items():
            if textures is None:
                raise ValueError("textured parameters need a texture list")
            keys = ["alpha_u", "alpha_v"] if key == "alpha" else [key]
            idx = None
            for k2 in keys:
                slot, colour = _TEX_SLOTS[k2]
                if idx is None:
                    idx = self._texture(tex, colour, textures)
                b.tex[slot] = idx
        if d["type"] == "disney":  # a textured parameter counts as given (ConfigHasValue)
            for key in textured:
                if key in F.DISNEY_FLAGS:
                    b.disney_flags |= F.DISNEY_FLAGS[key]
        return b

    def _bsdf_consts(self, d):
        b = F.BsdfDesc()
        t = d["type"]
        b.sample_visible = 1
        b.alpha_u = b.alpha_v = 0.1
        b.int_ior, b.ext_ior = 1.5046, 1.000277
        b.specular = 0.5
        b.specular_sampling_rate = b.clearcoat_sampling_rate = b.diffuse_sampling_rate = 1.0
        b.roughness = 0.5
        b.base_color[:] = [0.5, 0.5, 0.5]
        b.eta[:] = [0.1999899, 0.9220823, 1.0998751]
        b.k[:] = [3.9046354, 2.4476316, 2.1376500]
        b.reflectance[:] = [1, 1, 1]
        b.transmittance[:] = [1, 1, 1]
        if t == "diffuse":
            b.type = F.RAD_BSDF_DIFFUSE
            b.reflectance[:] = d.get("reflectance", [0.5] * 3)
        elif t in ("rough_metal", "rough_glass"):
            b.type = F.RAD_BSDF_ROUGH_METAL if t == "rough_metal" else F.RAD_BSDF_ROUGH_GLASS
            b.distribution = F.RAD_MICROFACET_BECKMANN if d.get("distribution") == "beckmann" else F.RAD_MICROFACET_GGX
            b.sample_visible = 1 if d.get("sample_visible", True) else 0
            a = d.get("alpha", 0.1)
            b.alpha_u = d.get("alpha_u", a)
            b.alpha_v = d.get("alpha_v", a)
            for key in ("reflectance", "transmittance", "eta", "k"):
                if key in d:
                    getattr(b, key)[:] = d[key]
            b.int_ior = d.get("int_ior", 1.5046)
            b.ext_ior = d.get("ext_ior", 1.000277)
        elif t == "disney":
            b.type = F.RAD_BSDF_DISNEY
            flags = 0
            for key, bit in F.DISNEY_FLAGS.items():
                if key == "eta":
                    continue
                if key in d and float(d[key]) != 0.0:
                    flags |= bit
            if "eta" in d:
                flags |= F.DISNEY_FLAGS["eta"]
                b.disney_eta = d["eta"]
            b.disney_flags = flags
            if "base_color" in d:
                b.base_color[:] = d["base_color"]
            for key in ("roughness", "anisotropic", "spec_trans", "sheen", "sheen_tint", "flatness", "spec_tint",
                        "metallic", "clearcoat", "clearcoat_gloss", "specular"):
                if key in d:
                    setattr(b, key, d[key])
            b.specular_sampling_rate = d.get("main_specular_sampling_rate", 1.0)
            b.clearcoat_sampling_rate = d.get("clearcoat_sampling_rate", 1.0)
            b.diffuse_sampling_rate = d.get("diffuse_reflectance_sampling_rate", 1.0)
        elif t in ("plastic", "rough_plastic"):
            b.type = F.RAD_BSDF_PLASTIC if t == "plastic" else F.RAD_BSDF_ROUGH_PLASTIC
            b.diffuse[:] = d.get("diffuse", [0.5] * 3)
            b.specular_reflectance[:] = d.get("specular", [1.0] * 3)
            b.int_ior = d.get("int_ior", 1.49)
            b.ext_ior = d.get("ext_ior", 1.000277)
            b.nonlinear = 1 if d.get("no_linear", False) else 0
            b.distribution = F.RAD_MICROFACET_BECKMANN if d.get("distribution") == "beckmann" else F.RAD_MICROFACET_GGX
            b.sample_visible = 1 if d.get("sample_visible", True) else 0
            b.alpha_u = b.alpha_v = d.get("alpha", 0.1)
        elif t == "perfect_mirror":
            b.type = F.RAD_BSDF_PERFECT_MIRROR
            b.reflectance[:] = d.get("reflectance", [1.0] * 3)
            b.eta[:] = d.get("eta", [0.0] * 3)
            b.k[:] = d.get("k", [1.0] * 3)
        elif t == "perfect_glass":
            b.type = F.RAD_BSDF_PERFECT_GLASS
            b.reflectance[:] = d.get("reflectance", [1.0] * 3)
            b.transmittance[:] = d.get("transmittance", [1.0] * 3)
            b.int_ior = d.get("int_ior", 1.5046)
            b.ext_ior = d.get("ext_ior", 1.000277)
        elif t in ("two_side", "mask"):
            b.type = F.RAD_BSDF_TWO_SIDE if t == "two_side" else F.RAD_BSDF_MASK
            b.opacity = d.get("opacity", 0.5)
            b.nested = -1  # set by bsdf_array
        else:
            raise ValueError(f"unsupported bsdf {t}")
        return b

    def bsdf_array(self, textures=None):
        """The BSDF descriptors, the wrappers' nested BSDFs appended after the
        builder's own entries (shape indices stay valid); textures collects
        the TextureDesc of textured parameters."""
        if textures is None:
            textures = []
        out = [self.bsdf_desc(d, textures) for d in self.bsdfs]

        def nest(i, d):
            inner = d.get("two_side") if d["type"] == "two_side" else d.get("nested")
            j = len(out)
            out.append(self.bsdf_desc(inner, textures))
            out[i].nested = j
            if inner["type"] in ("two_side", "mask"):
                nest(j, inner)

        for i, d in enumerate(self.bsdfs):
            if d["type"] in ("two_side", "mask"):
                nest(i, d)
        arr = (F.BsdfDesc * max(len(out), 1))()
        for i, b in enumerate(out):
            arr[i] = b
        return arr, len(out)

    def to_desc(self):
        self._keep = []
        n = len(self.shapes)
        shapes = (F.ShapeDesc * max(n, 1))()
        for i, s in enumerate(self.shapes):
            d = shapes[i]
            d.type = {"mesh": F.RAD_SHAPE_MESH, "sphere": F.RAD_SHAPE_SPHERE, "rectangle": F.RAD_SHAPE_RECTANGLE,
                      "cube": F.RAD_SHAPE_CUBE}[s.kind]
            d.bsdf, d.light = s.bsdf, s.light
            d.medium_inside, d.medium_outside = s.medium_inside, s.medium_outside
            d.flip_normal = 1 if s.flip_normal else 0
            d.to_world[:] = [float(x) for x in np.asarray(s.to_world, np.float32).reshape(-1)]
            if s.kind == "mesh":
                p, nn, uv, idx = s.mesh
                d.vertex_count = p.shape[0]
                d.triangle_count = idx.size // 3
                d.positions = self._ptr(p, C.c_float)
                d.normals = self._ptr(nn, C.c_float)
                d.uvs = self._ptr(uv, C.c_float)
                d.indices = self._ptr(idx, C.c_uint32)
            d.center[:] = s.center
            d.radius = s.radius
        tex_list = []
        bsdfs, bsdf_count = self.bsdf_array(tex_list)
        lights = (F.LightDesc * max(len(self.lights), 1))()
        for i, l in enumerate(self.lights):
            lights[i].type = {"area": F.RAD_LIGHT_AREA, "skybox": F.RAD_LIGHT_SKYBOX, "point": F.RAD_LIGHT_POINT,
                              "projection": F.RAD_LIGHT_PROJECTION}[l["type"]]
            lights[i].fov = l.get("fov", 90.0)
            lights[i].shape = l.get("shape", -1)
            if isinstance(l["radiance"], dict):  # a radiance (irradiance) texture; a projection keeps its scale
                lights[i].tex = self._texture(l["radiance"], True, tex_list)
                lights[i].radiance[:] = l["scale"] if l["type"] == "projection" else [1.0, 1.0, 1.0]
            else:
                lights[i].radiance[:] = l["radiance"]
            lights[i].to_world[:] = [float(x) for x in np.asarray(l.get("to_world", IDENTITY), np.float32).reshape(-1)]
        textures = (F.TextureDesc * max(len(tex_list), 1))()
        for i, td in enumerate(tex_list):
            textures[i] = td
        media = (F.MediumDesc * max(len(self.media), 1))()
        for i, m in enumerate(self.media):
            md = media[i]
            md.type = F.RAD_MEDIUM_HOMOGENEOUS if m["type"] == "homogeneous" else F.RAD_MEDIUM_HETEROGENEOUS
            ph = m["phase"]
            md.phase = F.RAD_PHASE_HG if ph["type"] == "hg_phase" else F.RAD_PHASE_ISOTROPIC
            md.g = ph.get("g", 0.8)
            md.sigma_t[:] = m.get("sigma_t", [1.0] * 3)
            md.albedo[:] = m["albedo"]
            md.scale = m["scale"]
            md.to_world[:] = [float(x) for x in np.asarray(m["to_world"], np.float32).reshape(-1)]
            if "grid" in m:
                g = m["grid"]
                md.grid_size[:] = [g.shape[2], g.shape[1], g.shape[0]]
                md.grid_channels = 1 if g.ndim == 3 else g.shape[3]
                md.grid_wrap = F.RAD_WRAP_CLAMP
                md.grid = self._ptr(g, C.c_float)
        cam = self.camera
        desc = F.SceneDesc()
        desc.shape_count = n
        desc.shapes = shapes
        desc.bsdf_count = bsdf_count
        desc.bsdfs = bsdfs
        desc.light_count = len(self.lights)
        desc.lights = lights
        desc.medium_count = len(self.media)
        desc.media = media
        desc.global_medium = self.global_medium
        desc.camera.to_world[:] = [float(x) for x in cam["to_world"].reshape(-1)]
        desc.camera.fov = cam["fov"]
        desc.camera.near_clip = cam["near"]
        desc.camera.far_clip = cam["far"]
        desc.camera.width = cam["width"]
        desc.camera.height = cam["height"]
        desc.texture_count = len(tex_list)
        desc.textures = textures
        self._keep += [shapes, bsdfs, lights, media, textures]
        return desc

    def render_desc(self, spp=None, seed=None, flags=0, path_slots=0, guiding=0, request_id=0):
        r = F.RenderDesc()
        r.integrator = F.RAD_INTEGRATOR_VOL_PATH if self.renderer["type"] == "vol_path" else F.RAD_INTEGRATOR_PATH
        r.max_depth = self.renderer["max_depth"]
        r.rr_depth = self.renderer["rr_depth"]
        r.spp = int(spp if spp is not None else self.sampler["sample_count"])
        r.seed = int(seed if seed is not None else self.sampler["seed"])
        r.flags = flags
        r.path_slots = path_slots
        r.guiding = guiding
        r.request_id = request_id
        return r

    # -- reference on-disk format ------------------------------------------------
    def write_json(self, directory, name="scene"):
        """Write `<name>.json` plus OBJ / .vol assets in the reference's format."""
        os.makedirs(directory, exist_ok=True)
        assets, entities = [], []
        fmt = lambda a: [float(x) for x in np.asarray(a, np.float32).reshape(-1)]  # noqa: E731
        for i, s in enumerate(self.shapes):
            e = {"to_world": fmt(s.to_world)}
            if s.kind == "mesh":
                fname = f"{s.asset}.obj"
                write_obj(os.path.join(directory, fname), *s.mesh)
                assets.append({"type": "model_obj", "name": s.asset, "location": fname})
                e["shape"] = {"type": "mesh", "asset_name": s.asset}
            elif s.kind == "sphere":
                e["shape"] = {"type": "sphere", "center": list(s.center), "radius": s.radius}
            elif s.kind == "rectangle":
                e["shape"] = {"type": "rectangle", "is_flip_normal": bool(s.flip_normal)}
            else:
                e["shape"] = {"type": "cube"}
            if s.bsdf >= 0 and not s.json_shape.get("auto_bsdf"):
                e["bsdf"] = self._json_value(self.bsdfs[s.bsdf], directory, assets)
            if s.light >= 0:
                l = self.lights[s.light]
                e["light"] = {"type": "area", "radiance": self._json_value(l["radiance"], directory, assets)}
            for key, idx in (("in_medium", s.medium_inside), ("out_medium", s.medium_outside)):
                if idx >= 0:
                    e[key] = self._medium_json(idx, directory, assets)
            entities.append(e)
        for l in self.lights:
            if l["type"] == "skybox":
                entities.append({"to_world": fmt(l["to_world"]),
                                 "light": {"type": "skybox", "radiance": self._json_value(l["radiance"], directory, assets)}})
            elif l["type"] == "point":
                entities.append({"light": {"type": "point", "position": l["position"], "intensity": l["radiance"]}})
            elif l["type"] == "projection":
                entities.append({"to_world": fmt(l["to_world"]),
                                 "light": {"type": "projection",
                                           "irradiance": self._json_value(l["irradiance"], directory, assets),
                                           "scale": l["scale"],
                                           "fov": l["fov"]}})
        cam = self.camera
        camj = {"type": "perspective", "fov": cam["fov"], "near": cam["near"], "far": cam["far"],
                "resolution": [cam["width"], cam["height"]], "to_world": fmt(cam["to_world"])}
        scene = {"camera": camj, "sampler": dict(self.sampler), "scene": entities, "assets": assets,
                 "renderer": dict(self.renderer)}
        if self.global_medium >= 0:
            scene["global_medium"] = self._medium_json(self.global_medium, directory, assets)
        path = os.path.join(directory, f"{name}.json")
        with open(path, "w") as f:
            json.dump(scene, f, indent=1)
        return path

    def _json_value(self, v, directory, assets):
        """A BSDF / texture value as JSON; bitmaps become 8-bit PNG image assets
        (texels must be multiples of 1/255; rows written top first, loaded
        with is_flip_y so that texel row 0 is the bottom row)."""
        if isinstance(v, dict) and v.get("type") == "bitmap":
            t = v["_texels"]
            u8 = np.rint(t * 255.0)
            if np.abs(u8 / 255.0 - t).max() > 1e-6 or u8.min() < 0 or u8.max() > 255:
                raise ValueError("bitmap texels must be k / 255 to be written as an 8-bit image")
            name = v.get("asset_name") or f"tex{len(assets)}"
            fname = f"{name}.png"
            write_png(os.path.join(directory, fname), u8.astype(np.uint8)[::-1])
            assets.append({"type": "image_default", "name": name, "location": fname, "channel": 3 if t.ndim == 3 else 1,
                           "is_to_linear": False, "is_flip_y": True})
            j = {"type": "bitmap", "asset_name": name, "filter": v["filter"], "wrap": v["wrap"]}
            if v["max_level"] is not None:
                j["max_level"] = int(v["max_level"])
            if v["anisotropic_level"] is not None:
                j["anisotropic_level"] = float(v["anisotropic_level"])
            return j
        if isinstance(v, dict) and v.get("type") == "chessboard":
            j = {"type": "chessboard", "remap": v["remap"],
                 "color_a": self._json_value(v["color_a"], directory, assets),
                 "color_b": self._json_value(v["color_b"], directory, assets)}
            if v["to_uv"] is not None:
                j["to_uv"] = [float(x) for x in v["to_uv"].reshape(-1)]
            return j
        if isinstance(v, dict):
            return {k: self._json_value(x, directory, assets) for k, x in v.items()}
        return v

    def _medium_json(self, idx, directory, assets):
        m = self.media[idx]
        j = {"type": m["type"], "albedo": m["albedo"], "scale": m["scale"], "phase": m["phase"]}
        if m["type"] == "homogeneous":
            j["sigma_t"] = m["sigma_t"]
        else:
            fname = f"{m['grid_name']}{idx}.vol"
            write_mitsuba_vol(os.path.join(directory, fname), m["grid"])
            aname = f"{m['grid_name']}{idx}"
            if not any(a["name"] == aname for a in assets):
                assets.append({"type": "volume_data_mitsuba_vol", "name": aname, "location": fname})
            j["sigma_t"] = {"type": "volume_grid", "asset_name": aname}
        return j


def write_png(path, u8):
    """8-bit grey (H x W) or RGB (H x W x 3) PNG, row 0 written first."""
    import struct
    import zlib
    u8 = np.ascontiguousarray(u8, np.uint8)
    h, w = u8.shape[:2]
    raw = b"".join(b"\x00" + u8[y].tobytes() for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 0 if u8.ndim == 2 else 2, 0, 0, 0))
                + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def write_obj(path, positions, normals, uvs, indices):
    """Wavefront OBJ with `f v/vt/vn` faces; %.9g round-trips float32 exactly."""
    p = np.asarray(positions, np.float32)
    lines = ["# generated"]
    lines += ["v %.9g %.9g %.9g" % tuple(r) for r in p.tolist()]
    if uvs is not None:
        lines += ["vt %.9g %.9g" % tuple(r) for r in np.asarray(uvs, np.float32).tolist()]
    if normals is not None:
        lines += ["vn %.9g %.9g %.9g" % tuple(r) for r in np.asarray(normals, np.float32).tolist()]
    idx = np.asarray(indices, np.int64).reshape(-1, 3) + 1
    if uvs is not None and normals is not None:
        lines += ["f %d/%d/%d %d/%d/%d %d/%d/%d" % (a, a, a, b, b, b, c, c, c) for a, b, c in idx.tolist()]
    elif normals is not None:
        lines += ["f %d//%d %d//%d %d//%d" % (a, a, b, b, c, c) for a, b, c in idx.tolist()]
    elif uvs is not None:
        lines += ["f %d/%d %d/%d %d/%d" % (a, a, b, b, c, c) for a, b, c in idx.tolist()]
    else:
        lines += ["f %d %d %d" % tuple(r) for r in idx.tolist()]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_mitsuba_vol(path, grid, bbox=(0, 0, 0, 1, 1, 1)):
    """Mitsuba VOL v3, float32, x fastest (core/src/volume_reader.cpp:65-124)."""
    g = np.ascontiguousarray(grid, np.float32)
    ch = 1 if g.ndim == 3 else g.shape[3]
    with open(path, "wb") as f:
        f.write(b"VOL" + bytes([3]))
        f.write(struct.pack("<iiiii", 1, g.shape[2], g.shape[1], g.shape[0], ch))
        f.write(struct.pack("<6f", *bbox))
        f.write(g.tobytes())


# --------------------------------------------------------------------------- benchmark scenes
def many_balls(seed=0, width=512, height=512, spp=16, n_spheres=64, slices=32, max_depth=8, rr_depth=3):
    """Plain path tracer, 'many balls': a 20x20 ground, n UV spheres of
    CreateSphere(r, slices) (~1k triangles each), Lambertian albedo U[0.1,0.9],
    one 2x2 area light at y=8 with radiance 10, pinhole camera, fov 45."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    ground = b.diffuse(0.5)
    b.rectangle(to_world=trs(angle=-90.0, axis=(1, 0, 0), scale=(10, 10, 10)), bsdf=ground)
    for i in range(n_spheres):
        c = rng.uniform([-5.0, 0.3, -5.0], [5.0, 3.0, 5.0])
        r = rng.uniform(0.2, 0.6)
        alb = rng.uniform(0.1, 0.9, 3)
        pos, nrm, uv, idx = create_sphere(r, slices)
        b.mesh(pos, idx, nrm, uv, to_world=trs(translate=c), bsdf=b.diffuse(alb), name=f"ball{i}")
    light = b.rectangle(to_world=trs(translate=(0, 8, 0), axis=(1, 0, 0), angle=90.0))
    b.area_light(light, 10.0)
    b.set_camera(width, height, 45.0, origin=(0.0, 7.0, -16.0), target=(0.0, 0.8, 0.0), far=1000.0)
    b.set_render("path", spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=seed)
    return b


def large_mesh(seed=1, width=1920, height=1080, spp=64, grid=2048, n_spheres=256, slices=46, max_depth=16,
               rr_depth=3):
    """Large-mesh stress scene: a grid x grid heightfield (6-octave fBm, amplitude
    2) plus n UV spheres (CreateSphere(r, 46), ~2k triangles) with materials from
    {GGX conductor alpha U[0.05,0.5], rough dielectric IOR 1.5, Disney}, a
    constant environment and a sun made of a small distant spherical emitter
    (the reference has no directional light type)."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    pos, nrm, uv, idx = create_grid(40.0, 40.0, grid, grid)
    pos[:, 1] = fbm_heightfield(grid, grid, 6, 2.0, seed).reshape(-1)
    # recompute smooth normals of the displaced field from central differences
    h = pos[:, 1].reshape(grid, grid)
    dz = np.gradient(h, axis=0) / (40.0 / (grid - 1))
    dx = np.gradient(h, axis=1) / (40.0 / (grid - 1))
    n = np.stack([-dx, np.ones_like(h), dz], -1).reshape(-1, 3)
    nrm = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    terrain_mat = b.diffuse((0.35, 0.3, 0.25))
    b.mesh(pos, idx, nrm, uv, bsdf=terrain_mat, name="terrain")
    hf = pos[:, 1].reshape(grid, grid)
    for i in range(n_spheres):
        cx, cz = rng.uniform(-18.0, 18.0, 2)
        r = rng.uniform(0.3, 1.0)
        gi = int(np.clip(round((20.0 - cz) / 40.0 * (grid - 1)), 0, grid - 1))
        gj = int(np.clip(round((cx + 20.0) / 40.0 * (grid - 1)), 0, grid - 1))
        cy = float(hf[gi, gj]) + r * 0.9
        kind = i % 3
        if kind == 0:
            m = b.rough_metal(alpha=rng.uniform(0.05, 0.5), reflectance=rng.uniform(0.6, 1.0, 3))
        elif kind == 1:
            m = b.rough_glass(alpha=rng.uniform(0.05, 0.3), int_ior=1.5, ext_ior=1.0)
        else:
            m = b.disney(base_color=rng.uniform(0.1, 0.9, 3), metallic=rng.uniform(0, 1),
                         roughness=rng.uniform(0.1, 0.9))
        p2, n2, uv2, i2 = create_sphere(r, slices)
        b.mesh(p2, i2, n2, uv2, to_world=trs(translate=(cx, cy, cz)), bsdf=m, name=f"obj{i}")
    sun = b.sphere(center=(300.0, 400.0, -200.0), radius=12.0)
    b.area_light(sun, 150.0)
    b.skybox((0.25, 0.32, 0.45))
    b.set_camera(width, height, 60.0, origin=(0.0, 9.0, -27.0), target=(0.0, 0.0, 0.0), far=5000.0)
    b.set_render("path", spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=seed)
    return b


def volume_box(seed=2, width=1024, height=1024, spp=64, res=256, max_depth=32, rr_depth=5, scale=5.0, albedo=0.8,
               g=0.3):
    """Heterogeneous medium in a 1x1x1 pass-through cube (res^3 4-octave Perlin
    density clamped >= 0, sigma_t scale 5, albedo 0.8, HG g=0.3) inside a
    Lambertian box room lit by one area light; rendered with `vol_path`."""
    b = SceneBuilder()
    dens = np.clip(perlin3d(res, 4, seed) * 2.0 + 0.3, 0.0, None).astype(np.float32)
    med = b.heterogeneous(dens, albedo=albedo, scale=scale, g=g, to_world=trs(scale=(0.5, 0.5, 0.5), translate=(0, 1.0, 0)))
    white, red, green = b.diffuse(0.7), b.diffuse((0.7, 0.15, 0.15)), b.diffuse((0.15, 0.6, 0.15))
    s = 2.5
    b.rectangle(to_world=trs(translate=(0, 0, 0), axis=(1, 0, 0), angle=-90, scale=(s, s, s)), bsdf=white)  # floor
    b.rectangle(to_world=trs(translate=(0, 5, 0), axis=(1, 0, 0), angle=90, scale=(s, s, s)), bsdf=white)  # ceiling
    b.rectangle(to_world=trs(translate=(0, 2.5, 2.5), axis=(0, 1, 0), angle=180, scale=(s, s, s)), bsdf=white)  # back
    b.rectangle(to_world=trs(translate=(-2.5, 2.5, 0), axis=(0, 1, 0), angle=90, scale=(s, s, s)), bsdf=red)
    b.rectangle(to_world=trs(translate=(2.5, 2.5, 0), axis=(0, 1, 0), angle=-90, scale=(s, s, s)), bsdf=green)
    b.cube(to_world=trs(translate=(0, 1.0, 0), scale=(0.5, 0.5, 0.5)), medium_inside=med)
    light = b.rectangle(to_world=trs(translate=(0, 4.98, 0), axis=(1, 0, 0), angle=90, scale=(0.7, 0.7, 0.7)))
    b.area_light(light, 12.0)
    b.set_camera(width, height, 50.0, origin=(0.0, 2.5, -8.5), target=(0.0, 1.8, 0.0), far=1000.0)
    b.set_render("vol_path", spp=spp, max_depth=max_depth, rr_depth=rr_depth, seed=seed)
    return b


def guiding_room(seed=3, width=1280, height=720, spp=256, n_boxes=32, max_depth=-1, light_x=9.0, light_size=1.0,
                 light_radiance=32000.0):
    """Difficult indirect lighting: a 6x3x6 closed room lit only through a
    0.5x0.5 gap in the x = 3 wall by an exterior area light (light_size
    square, centred at x = light_x, facing the gap), with n glossy GGX cubes.
    The default light, 1x1 at 6 m behind the gap, throws a beam about 2 m wide
    onto the far wall: that spot lights the rest of the room, which BSDF
    sampling finds rarely."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    wall = b.diffuse(0.65)
    # floor / ceiling / walls facing inward (rectangles are [-1,1]^2 in local xy)
    b.rectangle(to_world=trs(translate=(0, 0, 0), axis=(1, 0, 0), angle=-90, scale=(3, 3, 1)), bsdf=wall)
    b.rectangle(to_world=trs(translate=(0, 3, 0), axis=(1, 0, 0), angle=90, scale=(3, 3, 1)), bsdf=wall)
    b.rectangle(to_world=trs(translate=(-3, 1.5, 0), axis=(0, 1, 0), angle=90, scale=(3, 1.5, 1)), bsdf=wall)
    b.rectangle(to_world=trs(translate=(0, 1.5, -3), axis=(0, 1, 0), angle=0, scale=(3, 1.5, 1)), bsdf=wall)
    b.rectangle(to_world=trs(translate=(0, 1.5, 3), axis=(0, 1, 0), angle=180, scale=(3, 1.5, 1)), bsdf=wall)
    # +x wall with a 0.5 x 0.5 gap centred at (y=1.5, z=0): four panels
    rot = dict(axis=(0, 1, 0), angle=-90)
    b.rectangle(to_world=trs(translate=(3, 0.625, 0), scale=(3, 0.625, 1), **rot), bsdf=wall)
    b.rectangle(to_world=trs(translate=(3, 2.375, 0), scale=(3, 0.625, 1), **rot), bsdf=wall)
    b.rectangle(to_world=trs(translate=(3, 1.5, -1.625), scale=(1.375, 0.25, 1), **rot), bsdf=wall)
    b.rectangle(to_world=trs(translate=(3, 1.5, 1.625), scale=(1.375, 0.25, 1), **rot), bsdf=wall)
    for _ in range(n_boxes):
        sx, sy, sz = rng.uniform(0.1, 0.35, 3)
        cx, cz = rng.uniform(-2.5, 2.2, 2)
        m = b.rough_metal(alpha=rng.uniform(0.08, 0.3), reflectance=rng.uniform(0.5, 0.95, 3))
        b.cube(to_world=trs(translate=(cx, sy, cz), axis=(0, 1, 0), angle=rng.uniform(0, 90), scale=(sx, sy, sz)),
               bsdf=m)
    h = 0.5 * light_size
    light = b.rectangle(to_world=trs(translate=(light_x, 1.5, 0), axis=(0, 1, 0), angle=-90, scale=(h, h, 1.0)))
    b.area_light(light, light_radiance)
    b.set_camera(width, height, 70.0, origin=(-2.6, 1.6, -2.6), target=(1.0, 1.0, 1.0), far=1000.0)
    b.set_render("path", spp=spp, max_depth=max_depth, rr_depth=3, seed=seed)
    return b
