"""The C-ABI library loads, exports every function include/rad_hip.h declares,
and the ctypes mirrors match the C struct sizes.  No device calls."""
import ctypes as C
import os
import re

import pytest

from radiantiumplusplus_amd import _ffi as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rad_hip.h")
LIB = os.path.join(ROOT, "radiantiumplusplus_amd", "libradhip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(rad_[a-z0-9_]+)\s*\(", text)
    return sorted(set(n for n in names if not n.startswith("rad_status")))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first"
    lib = C.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"undefined entry points: {missing}"


@pytest.mark.parametrize("name", sorted(F.ABI_STRUCTS))
def test_struct_layout_matches(name):
    lib = C.CDLL(LIB)
    lib.rad_abi_sizeof.restype = C.c_size_t
    lib.rad_abi_sizeof.argtypes = [C.c_char_p]
    assert lib.rad_abi_sizeof(name.encode()) == C.sizeof(F.ABI_STRUCTS[name])


def test_init_without_device_fails_cleanly():
    """On a host without a GPU the library reports an error instead of
    falling back to any CPU path."""
    import subprocess
    import sys

    code = (
        "import ctypes as C, sys\n"
        f"lib = C.CDLL({LIB!r})\n"
        "h = C.c_void_p()\n"
        "lib.rad_hip_init.argtypes=[C.c_int32, C.POINTER(C.c_void_p)]\n"
        "st = lib.rad_hip_init(0, C.byref(h))\n"
        "lib.rad_hip_last_error.restype = C.c_char_p\n"
        "print(st, lib.rad_hip_last_error().decode())\n"
    )
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = "-1"  # hide any device
    env["ROCR_VISIBLE_DEVICES"] = "-1"
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    status = int(r.stdout.split()[0])
    assert status != 0
