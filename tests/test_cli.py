"""The command-line renderer (counterpart of rad.offline.cli/main.cpp:8-71):
argument errors on the CPU; a JSON scene rendered end to end on the GPU,
written as EXR, identical to the same render through the library API."""
import os
import subprocess

import numpy as np
import pytest

from radiantiumplusplus_amd import hip, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "radiantiumplusplus_amd", "bin", "rad_offline_cli")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)


def test_cli_requires_a_scene():
    r = run()
    assert r.returncode == 2 and "--scene" in r.stderr


def test_cli_reports_a_missing_file(tmp_path):
    r = run("--scene", str(tmp_path / "none.json"))
    assert r.returncode == 2 and "cannot open file" in r.stderr


def read_exr(path, w, h):
    raw = open(path, "rb").read()
    data = raw[-h * (8 + w * 12):]
    img = np.zeros((h, w, 3), np.float32)
    for y in range(h):
        block = data[y * (8 + w * 12) + 8:(y + 1) * (8 + w * 12)]
        planes = np.frombuffer(block, np.float32).reshape(3, w)  # B, G, R
        img[y] = planes[::-1].T
    return img


@pytest.mark.gpu
def test_cli_renders_json_scene(tmp_path, hip_ctx):
    b = scenes.many_balls(seed=1, width=40, height=30, spp=16, n_spheres=8)
    path = b.write_json(str(tmp_path), "balls")
    r = run("--scene", path)
    assert r.returncode == 0, r.stderr
    assert "render used time" in r.stdout
    img = read_exr(str(tmp_path / "balls.exr"), 40, 30)
    s = hip.Scene.from_builder(hip_ctx, b)
    ref = s.render(b.render_desc())[0]
    assert np.array_equal(img, ref)
