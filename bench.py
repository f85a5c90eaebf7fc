#!/usr/bin/env python3
"""Throughput of the gpuPathMIS renderer on one MI355X.

Workload (one step = one full render): displaced icosphere of subdivision 8
(1,310,720 triangles, radii 1 + 0.08 fBm), GGX rough gold conductor
(alpha 0.3), diffuse ground, synthetic 512 x 256 environment map sampled by
its CDF, pinhole camera, 1024 x 768, 256 samples per pixel, Gaussian filter.

Prints one JSON line: path samples/s (whole job), p50/p99 step latency,
rays/s, the HBM roofline of the dominant kernel (k_extend, timed with HIP
events on the context stream) and the oracle's CPU throughput on a bounded
sample of the same workload.
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
# Algorithmic bytes (FP32 SoA queues, excluding BVH node traffic):
#   extension segment 296 B, shadow ray 190 B, path 128 B (whole render);
#   the extension kernel's own share of a segment is the ray read (32 B), the
#   hit write (16 B) and the triangle fetch (40 B) = 88 B per ray.
SEG_B, SHADOW_B, PATH_B, EXTEND_B = 296, 190, 128, 88


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--alpha", type=float, default=0.3)
    ap.add_argument("--max-paths", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    return ap.parse_args()


def cpu_baseline(scene, target_s):
    """Oracle (C++ restatement, the 'port') on the box's host cores, bounded sample."""
    from oracle import pyoracle as O
    threads = min(16, os.cpu_count() or 1)
    _, _, st = O.render(scene, req_seed=0, spp=(0, 1), threads=threads)
    t1 = float(st[3]) * 1e-6
    paths, secs, spp_done = int(st[2]), t1, 1
    extra = int(max(0, min(64, round(target_s / max(t1, 1e-3)) - 1)))
    if extra > 0:
        _, _, st = O.render(scene, req_seed=0, spp=(1, 1 + extra), threads=threads)
        paths += int(st[2])
        secs += float(st[3]) * 1e-6
        spp_done += extra
    return {"value": paths / secs, "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": f"same scene, all {scene.width}x{scene.height} pixels at {spp_done} spp "
                      f"({paths} paths, {secs:.1f} s render time, BVH build excluded)"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU sharding is described in DESIGN.md and not run", file=sys.stderr)
        sys.exit(2)
    import hikari_amd as hk
    from hikari_amd import scenes

    t = time.time()
    scene = scenes.displaced_sphere(alpha=a.alpha)
    print(f"# scene built in {time.time() - t:.1f} s: {scene.n_triangles} triangles", file=sys.stderr, flush=True)
    ctx = hk.Context(0)
    ctx.set_option("profile", 1)
    if a.max_paths:
        ctx.set_option("max_paths", a.max_paths)
    t = time.time()
    gs = ctx.scene(scene)
    print(f"# scene resident in HBM ({gs.device_bytes() / 2**20:.0f} MiB) after {time.time() - t:.1f} s",
          file=sys.stderr, flush=True)
    for i in range(a.warmup):
        _, st = gs.render(a.spp, seed=1000 + i)
        print(f"# warmup {i}: {st['gpu_ms']:.1f} ms", file=sys.stderr, flush=True)
    lat, tot = [], {"paths": 0, "extension_rays": 0, "shadow_rays": 0, "trace_ms": 0.0, "trace_launches": 0,
                    "gpu_ms": 0.0, "node_visits": 0, "triangle_tests": 0}
    t_all = time.perf_counter()
    for i in range(a.steps):
        t0 = time.perf_counter()
        img, st = gs.render(a.spp, seed=i)
        lat.append(time.perf_counter() - t0)
        for k in tot:
            tot[k] += st[k]
        print(f"# step {i}: {lat[-1] * 1e3:.1f} ms, {st['paths'] / lat[-1]:.3e} paths/s", file=sys.stderr,
              flush=True)
    wall = time.perf_counter() - t_all
    value = tot["paths"] / wall
    rays = tot["extension_rays"] + tot["shadow_rays"]
    ext_avg_ms = tot["trace_ms"] / max(1, tot["trace_launches"])
    ext_bytes_per_launch = EXTEND_B * tot["extension_rays"] / max(1, tot["trace_launches"])
    achieved = ext_bytes_per_launch / (ext_avg_ms * 1e-3) / 1e9 if ext_avg_ms > 0 else 0.0
    render_bytes = SEG_B * tot["extension_rays"] + SHADOW_B * tot["shadow_rays"] + PATH_B * tot["paths"]
    traffic = None
    tp = os.path.join(ROOT, "profiles", "k_extend_traffic.json")
    if os.path.exists(tp):
        with open(tp) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    out = {
        "metric": "path samples/sec at matched relMSE vs plain GPU path; p50/p99 render latency",
        "value": value, "unit": "path samples/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "displaced icosphere-8 (1,310,720 tris), GGX rough conductor alpha "
                               f"{a.alpha}, env map 512x256, 1024x768, {a.spp} spp, plain gpuPathMIS",
                   "spp": a.spp, "width": scene.width, "height": scene.height, "triangles": scene.n_triangles},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
        "rays_per_s": rays / wall,
        "rays_per_path": rays / max(1, tot["paths"]),
        "roofline": {"bound": "hbm", "kernel": "k_extend", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "launch_ms": ext_avg_ms, "rays_per_launch": tot["extension_rays"] / max(1, tot["trace_launches"])},
        "render_roofline_frac": render_bytes / wall / (HBM_PEAK_GBS * 1e9),
    }
    if tot["node_visits"]:  # builds with traversal counters only
        out["traversal"] = {"node_visits_per_ray": tot["node_visits"] / max(1, rays),
                            "triangle_tests_per_ray": tot["triangle_tests"] / max(1, rays)}
    if not a.no_cpu_baseline:
        t = time.time()
        out["cpu_baseline"] = cpu_baseline(scene, a.cpu_seconds)
        print(f"# cpu baseline in {time.time() - t:.1f} s", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
