"""MI355X-native `path` / `vol_path` integrators of RadiantiumPlusPlus.

Native code lives in libradhip.so (HIP kernels for gfx950 behind the C ABI of
include/rad_hip.h); `hip` is its Python front end, `scenes` builds scene
descriptions and writes the reference's JSON / OBJ / .vol formats.
"""
from . import _ffi  # noqa: F401

__all__ = ["hip", "scenes"]
