#!/usr/bin/env python3
"""Benchmark of the GPU `path` integrator on the large-mesh stress scene.

Workload (default): 1920x1080, 64 spp, max depth 16, a 2048x2048-vertex fBm
terrain (8.38M triangles) plus 256 UV spheres of 2,116 triangles with GGX
conductor / rough dielectric / Disney materials, a constant environment and
a small distant spherical "sun" emitter.  One step = one full render of that
frame.  Prints one JSON line (see README of the measurement contract).

The CPU baseline is the oracle (the CPU restatement of the reference's
integrator, its own binned-SAH BVH, OpenMP over tiles) timed on a bounded
window of the same frame at full spp.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRACE_BYTES_PER_RAY = 52  # ray read 32 B + hit write 20 B per closest-hit query


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--config", default="large_mesh", choices=["large_mesh", "many_balls", "volume_box"])
    p.add_argument("--spp", type=int, default=None)
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample duration")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--grid", type=int, default=2048)
    return p.parse_args()


def make_scene(cfg, args):
    from radiantiumplusplus_amd import scenes

    if cfg == "large_mesh":
        b = scenes.large_mesh(grid=args.grid)
        name = f"large_mesh_terrain{args.grid}_256spheres_1080p_64spp"
    elif cfg == "many_balls":
        b = scenes.many_balls()
        name = "many_balls_64spheres_512x512_16spp"
    else:
        b = scenes.volume_box()
        name = "volume_box_perlin256_1024x1024_64spp"
    if args.spp:
        b.sampler["sample_count"] = args.spp
    return b, name


def cpu_baseline(b, seconds):
    """Time the oracle on a centred window of the frame at full spp."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    threads = os.cpu_count() or 1
    t0 = time.time()
    o = O.OracleScene(b)
    build_s = time.time() - t0
    W, H = b.camera["width"], b.camera["height"]
    spp = b.sampler["sample_count"]
    # calibrate with a small window, then size the measured window to ~seconds
    side = 16
    win = (W // 2 - side // 2, W // 2 + side // 2, H // 2 - side // 2, H // 2 + side // 2)
    _, _, t = o.render(window=win, threads=threads, want_sumsq=False)
    rate = side * side * spp / max(t, 1e-6)
    npx = max(64, int(rate * seconds / spp))
    side = max(8, int(np.sqrt(npx)))
    side = min(side, min(W, H))
    win = (W // 2 - side // 2, W // 2 - side // 2 + side, H // 2 - side // 2, H // 2 - side // 2 + side)
    _, _, t = o.render(window=win, threads=threads, want_sumsq=False)
    samples = side * side * spp
    return {
        "value": samples / t,
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle path integrator, centred {side}x{side}-pixel window at {spp} spp "
                  f"({samples} samples, {t:.1f} s; BVH build {build_s:.1f} s excluded)",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from radiantiumplusplus_amd import _ffi as F
    from radiantiumplusplus_amd import hip

    t0 = time.time()
    b, name = make_scene(args.config, args)
    gen_s = time.time() - t0
    ctx = hip.Context(0)
    t0 = time.time()
    scene = hip.Scene.from_builder(ctx, b)
    setup_s = time.time() - t0
    info = scene.info()
    W, H = b.camera["width"], b.camera["height"]
    spp = b.sampler["sample_count"]
    rd = b.render_desc(spp=spp, seed=1, flags=F.RAD_RENDER_PROFILE)
    print(f"# scene {name}: {info.primitive_count} primitives, {info.node_count} BVH nodes, "
          f"build {info.bvh_build_ms:.0f} ms, generation {gen_s:.1f} s, upload+build {setup_s:.1f} s",
          file=sys.stderr, flush=True)
    for i in range(args.warmup):
        _, _, st = scene.render(rd)
        print(f"# warmup {i}: {st.render_ms:.1f} ms", file=sys.stderr, flush=True)
    times, stats = [], []
    t_all = time.time()
    for i in range(args.steps):
        rd.seed = 100 + i
        t = time.time()
        _, _, st = scene.render(rd)
        times.append(time.time() - t)
        stats.append(st)
        print(f"# step {i}: {times[-1] * 1e3:.1f} ms, {st.segments} segments, {st.iterations} iterations",
              file=sys.stderr, flush=True)
    total = time.time() - t_all
    samples = W * H * spp * args.steps
    ms_step = 1e3 * total / args.steps
    ext_ms = sum(s.kernel_ms[F.RAD_KERNEL_EXTEND] for s in stats)
    ext_launch = sum(s.kernel_launches[F.RAD_KERNEL_EXTEND] for s in stats)
    segs = sum(s.segments for s in stats)
    achieved = TRACE_BYTES_PER_RAY * segs / (ext_ms * 1e-3) / 1e9 if ext_ms > 0 else 0.0
    out = {
        "metric": "samples/s (GPU wavefront path tracer)",
        "value": samples / total,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": name,
            "width": W,
            "height": H,
            "spp": spp,
            "max_depth": b.renderer["max_depth"],
            "primitives": int(info.primitive_count),
            "integrator": b.renderer["type"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_extend (closest-hit traversal)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_ms": ext_ms / max(ext_launch, 1),
            "rays_per_s": segs / (ext_ms * 1e-3) if ext_ms > 0 else 0.0,
        },
        "stats": {
            "segments_per_step": segs / args.steps,
            "shadow_rays_per_step": sum(s.shadow_rays for s in stats) / args.steps,
            "iterations_per_step": sum(s.iterations for s in stats) / args.steps,
            "slots": stats[0].slots,
            "kernel_ms_per_step": {k: sum(s.kernel_ms[i] for s in stats) / args.steps
                                   for i, k in enumerate(["regen", "extend", "shade", "shadow"])},
            "bvh_build_ms": info.bvh_build_ms,
        },
    }
    if not args.no_cpu:
        cb = cpu_baseline(b, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
