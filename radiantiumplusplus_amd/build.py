"""Build recipe for the native library (libradhip.so) and the CLI.

Every source is compiled with hipcc for gfx950 in parallel, then linked into
radiantiumplusplus_amd/libradhip.so (in-tree, so it travels with the
repository snapshot).  `build(force=True)` always recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shlex
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libradhip.so")
CLI = os.path.join(PKG, "bin", "rad_offline_cli")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RAD_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wall", "-Wno-unused-variable", "-Wno-unused-function", "-Wno-unused-but-set-variable",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{CSRC}",
]

SOURCES = [
    "bvh_build.cpp",
    "bvh_gpu.hip",
    "scene_host.cpp",
    "render_job.cpp",
    "capi.cpp",
    "path_kernels.hip",
    "volpath_kernels.hip",
    "texture_kernels.hip",
    "guide_kernels.hip",
    "server.cpp",
    "host/json.cpp",
    "host/loader.cpp",
    "host/image.cpp",
    "host/exr.cpp",
]
CLI_SOURCES = ["host/cli.cpp", "host/renderer.cpp"]


def _sources():
    out = []
    for s in SOURCES:
        p = os.path.join(CSRC, s)
        if os.path.exists(p):
            out.append(p)
    return out


def _obj_for(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    return os.path.join(OBJ, rel + ".o")


def _compile(src, force):
    obj = _obj_for(src)
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= _newest_dep(src):
        return obj, ""
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC] + COMMON + lang + ["-c", src, "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(shlex.quote(c) for c in cmd)}\n{r.stderr[-6000:]}")
    os.replace(obj + ".tmp", obj)
    return obj, r.stderr


def _newest_dep(src):
    newest = os.path.getmtime(src)
    for d, _, files in os.walk(CSRC):
        for f in files:
            if f.endswith(".h"):
                newest = max(newest, os.path.getmtime(os.path.join(d, f)))
    newest = max(newest, os.path.getmtime(os.path.join(ROOT, "include", "rad_hip.h")))
    return newest


def build(force=True, jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for obj, log in ex.map(lambda s: _compile(s, force), srcs):
            objs.append(obj)
            if verbose and log.strip():
                print(log, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    # command-line front end (mirror of rad.offline.cli)
    cli_srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cli_srcs:
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        cmd = ["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", "-o", CLI + ".tmp"] + cli_srcs + [
            f"-L{PKG}", "-lradhip", f"-Wl,-rpath,$ORIGIN/..", "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI build failed\n{r.stderr[-4000:]}")
        os.replace(CLI + ".tmp", CLI)
    return LIB


def build_oracle(force=True):
    """Compile the CPU restatement used as the checker (oracle/Makefile)."""
    d = os.path.join(ROOT, "oracle")
    args = ["make", "-C", d, "-j4"]
    if force:
        args.insert(1, "-B")
    r = subprocess.run(args, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stderr[-4000:]}")
    return os.path.join(d, "librefcpu.so")


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose=True))
