#!/usr/bin/env python3
"""Throughput and latency of the gpuPathMIS renderer on one MI355X.

Default run (one JSON line):
  * value: the plain single-request path on the displaced-sphere workload
    (one step = one full render): icosphere of subdivision 8 (1,310,720
    triangles, radii 1 + 0.08 fBm), GGX rough gold conductor (alpha 0.3),
    diffuse ground, synthetic 512 x 256 environment map sampled by its CDF,
    pinhole camera, 1024 x 768, 256 samples per pixel, Gaussian filter; with
    p50/p99 step latency, rays/s, the HBM roofline of the dominant kernel
    (k_extend, timed with HIP events on the context stream) and the relMSE of
    a reduced film of the same scene against the oracle at equal streams;
  * batch: the 64-request mix over the same scene (res {64..1024}^2, spp
    {16..1024}, seeded poses on a radius-4 sphere) in one hk_render_batch
    against the same requests rendered one at a time with hk_render, images
    compared bit for bit, with the plain path's rate per film size;
  * serve: Poisson arrivals over three resident scenes (box, displaced
    sphere, full heightfield) through hk_server, with its express lanes (the
    default) and without them (option "express" = 0), against a first-come-
    first-served queue of plain hk_render calls on the same trace: p50/p99
    latency and path samples/s of each, at --qps (10: the GPU ~70% busy) and
    for the same requests at three times that rate (serve_saturated);
  * c0: the 256 x 256 box at 64 spp (the configuration the reference
    renderer was timed on, 3.35e5 path samples/s on 8 host threads) as plain
    renders, beside that rate;
  * cpu_baseline: the oracle (the C++ restatement) on the host cores on a
    bounded sample of the displaced-sphere workload, with its measured ratio
    to the reference renderer on the box scene.

--workload c2|batch|serve|accel|copy runs one leg alone (its own JSON line).
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
METRIC = "path samples/sec at matched relMSE vs plain GPU path; p50/p99 render latency"

# The oracle against the reference renderer, both on the 256 x 256 box
# scene at 64 spp on the same 8 host cores: the reference's pathMIS (built
# headless with its TBB loop replaced by a std::thread driver) measured
# 3.35e5 path samples/s and 26.3 rays per path; the oracle measured
# 3.74e5-3.86e5 and 24.3 rays per path.
ANCHOR_C0 = {"scene": "256x256 box, 64 spp, 8 threads of one host", "reference_path_samples_per_s": 3.35e5,
             "reference_rays_per_path": 26.3, "oracle_path_samples_per_s": 3.80e5, "oracle_rays_per_path": 24.3,
             "oracle_over_reference": 3.80e5 / 3.35e5}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--alpha", type=float, default=0.3)
    ap.add_argument("--max-paths", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0, help="k_extend refill threshold, 1..64 (0: library default)")
    ap.add_argument("--refill-shadow", type=int, default=0, help="k_shadow refill threshold (0: library default)")
    ap.add_argument("--fetch-batch", type=int, default=0,
                    help="queue entries a persistent wave reserves per atomic, in 64s (0: library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="default run: skip the batch leg")
    ap.add_argument("--no-serve", action="store_true", help="default run: skip the serving leg")
    ap.add_argument("--graph", type=int, default=1, help="1: replay bounces from captured HIP graphs")
    ap.add_argument("--autotune", type=int, default=0,
                    help="1: pick max_paths / grid / tile with hk_autotune before timing (results do not change); "
                         "the defaults are its usual choice on these scenes")
    ap.add_argument("--bvh-builder", choices=["host", "device"], default="host",
                    help="BVH construction: binned SAH on the host or linear BVH on the device")
    ap.add_argument("--bvh-cell-bits", type=int, default=-1,
                    help="device BVH: Morton bits of the cells under the SAH upper tree (-1: library default)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", choices=["c1", "c2", "batch", "serve", "accel", "copy"], default="c1")
    ap.add_argument("--requests", type=int, default=0, help="batch/serve: number of requests (default 64 / 40)")
    ap.add_argument("--qps", type=float, default=10.0, help="serve: Poisson arrival rate")
    ap.add_argument("--budget-gb", type=float, default=160.0, help="serve: HBM budget of the server")
    ap.add_argument("--schedule", type=int, default=1, help="batch: 1 fewest samples left first (library "
                    "default), 0 first come first served")
    ap.add_argument("--detail", default="", help="batch / serve: write per-request timings (JSON) to this file")
    return ap.parse_args()


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


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
    value = paths / secs
    return {"value": value, "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": f"same scene, all {scene.width}x{scene.height} pixels at {spp_done} spp "
                      f"({paths} paths, {secs:.1f} s render time, BVH build excluded)",
            "anchor_c0": ANCHOR_C0,
            "reference_estimate": value / ANCHOR_C0["oracle_over_reference"]}


def c0_leg(gs_box, spp=64, renders=5):
    """The box scene (256 x 256, 64 spp: the configuration the reference
    renderer was timed on, ANCHOR_C0) as plain renders on the GPU, beside
    the reference's measured CPU rate."""
    gs_box.render(spp, seed=900)
    paths = rays = 0
    t0 = time.perf_counter()
    for i in range(renders):
        _, st = gs_box.render(spp, seed=i)
        paths += st["paths"]
        rays += st["extension_rays"] + st["shadow_rays"]
    wall = time.perf_counter() - t0
    value = paths / wall
    return {"workload": f"256x256 box, {spp} spp, plain gpuPathMIS (one render per step)", "value": value,
            "unit": "path samples/s", "ms_per_render": wall / renders * 1e3, "rays_per_path": rays / max(1, paths),
            "reference_cpu": {"value": ANCHOR_C0["reference_path_samples_per_s"], "threads": 8,
                              "rays_per_path": ANCHOR_C0["reference_rays_per_path"]},
            "gpu_over_reference_cpu": value / ANCHOR_C0["reference_path_samples_per_s"]}


def _pct(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q))


def batch_leg(gs, n_requests, schedule, detail=""):
    """The 64-request mix in one hk_render_batch vs one hk_render each."""
    import hikari_amd as hk  # noqa: F401
    from hikari_amd import workloads
    reqs = workloads.batch_requests(n=n_requests)
    total = sum(r.width * r.height * r.spp for r in reqs)
    log(f"batch: {len(reqs)} requests, {total / 1e9:.2f} G path samples")
    gs.ctx.set_option("schedule", schedule)
    t0 = time.perf_counter()
    imgs_b, st = gs.render_batch(reqs)
    t_batch = time.perf_counter() - t0
    log(f"batch: batched {t_batch:.2f} s")
    lat_plain, same, rows = [], 0, []
    t0 = time.perf_counter()
    for r, ib in zip(reqs, imgs_b):
        t1 = time.perf_counter()
        img, pst = gs.render(0, request=r)
        lat_plain.append(time.perf_counter() - t0)
        rows.append({"pixels": r.width * r.height, "spp": r.spp, "alone_ms": (time.perf_counter() - t1) * 1e3,
                     "gpu_ms": pst["gpu_ms"], "bounces": pst["bounces_launched"]})
        same += int(np.array_equal(img.view(np.uint32), ib.view(np.uint32)))
    t_plain = time.perf_counter() - t0
    gs.ctx.set_option("schedule", 1)
    log(f"batch: one at a time {t_plain:.2f} s, identical images {same}/{len(reqs)}")
    if detail:
        with open(detail, "w") as fh:
            json.dump(rows, fh)
    # plain rate per film size, and the batch ratio it allows: a batch can at
    # best run every sample at the largest films' plain rate
    by = {}
    for row in rows:
        s = by.setdefault(row["pixels"], [0, 0.0])
        s[0] += row["pixels"] * row["spp"]
        s[1] += row["alone_ms"] * 1e-3
    rate = {str(int(round(k ** 0.5))): v[0] / v[1] for k, v in sorted(by.items())}
    best = max(rate.values())
    return {"value": total / t_batch, "unit": "path samples/s", "requests": len(reqs), "path_samples": total,
            "batched_s": t_batch, "plain_s": t_plain, "plain_value": total / t_plain,
            "batch_over_plain": t_plain / t_batch, "identical_images": same,
            "plain_rate_by_film_edge": rate, "bound_at_best_plain_rate": t_plain / (total / best),
            "rays_per_s": (st["extension_rays"] + st["shadow_rays"]) / t_batch,
            "schedule": ["fifo", "shortest_first"][schedule]}


def _run_server(ctx, gss, trace, budget_gb, express):
    """The trace through hk_server at its arrival times; per-request rows."""
    import threading
    import hikari_amd as hk
    ctx.set_option("express", express)
    srv = hk.Server(ctx, hbm_budget=int(budget_gb * 2 ** 30))
    for g in gss:
        srv.add_scene(g)
    ctx.set_option("express", 1)
    tickets = [None] * len(trace)
    t_start = time.perf_counter()

    def feeder():
        for i, (arr, sid, r) in enumerate(trace):
            d = t_start + arr - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(sid, r)

    th = threading.Thread(target=feeder)
    th.start()
    lat, t_last, rows = [], 0.0, []
    for i in range(len(trace)):
        while tickets[i] is None:
            time.sleep(0.001)
        _, res = srv.wait(tickets[i])
        lat.append(res["latency_ms"] * 1e-3)
        t_last = max(t_last, time.perf_counter() - t_start)
        arr, sid, r = trace[i]
        rows.append({"scene": sid, "pixels": r.width * r.height, "spp": r.spp, "arrival_s": arr, "lane": res["lane"],
                     "server_ms": res["latency_ms"], "queue_ms": res["queue_ms"], "submit_ms": res["submit_ms"]})
    th.join()
    stats = srv.stats()
    srv.close()
    return lat, t_last - trace[0][0], stats, rows


def serve_leg(ctx, gss, trace, budget_gb, detail="", service=None):
    """Poisson arrivals through hk_server, with and without its express
    lanes, vs a FCFS queue of plain renders.  service: the requests' render
    times alone, when already measured for the same requests."""
    total = sum(r.width * r.height * r.spp for _, _, r in trace)
    # plain: each request alone; first come first served on one queue
    if service is None:
        service = []
        for _, sid, r in trace:
            t0 = time.perf_counter()
            gss[sid].render(0, request=r)
            service.append(time.perf_counter() - t0)
    t_free, lat_plain = 0.0, []
    for (arr, _, _), sv in zip(trace, service):
        t_free = max(arr, t_free) + sv
        lat_plain.append(t_free - arr)
    span_plain = t_free - trace[0][0]
    log(f"serve: plain FCFS busy {sum(service):.2f} s, span {span_plain:.2f} s")
    out = {}
    for express in (1, 0):
        lat, span, stats, rows = _run_server(ctx, gss, trace, budget_gb, express)
        for row, sv, lp in zip(rows, service, lat_plain):
            row["alone_ms"], row["plain_fcfs_ms"] = sv * 1e3, lp * 1e3
        log(f"serve: server (express lanes {'on' if express else 'off'}) span {span:.2f} s, "
            f"p50 {_pct(lat, 50) * 1e3:.0f} ms, p99 {_pct(lat, 99) * 1e3:.0f} ms "
            f"(plain FCFS p50 {_pct(lat_plain, 50) * 1e3:.0f} ms, p99 {_pct(lat_plain, 99) * 1e3:.0f} ms)")
        out[express] = {"value": total / span, "span_s": span,
                        "latency_ms": {"p50": _pct(lat, 50) * 1e3, "p99": _pct(lat, 99) * 1e3},
                        "served_over_plain": span_plain / span,
                        "express_requests": sum(1 for x in rows if x["lane"] == 1),
                        "server": {"bounces": stats["bounces"], "shared_bounces": stats["shared_bounces"],
                                   "max_in_flight": stats["max_in_flight"], "fixed_gb": stats["fixed_bytes"] / 2 ** 30,
                                   "hbm_budget_gb": budget_gb},
                        "rows": rows}
    if detail:
        with open(detail, "w") as fh:
            json.dump({"express": out[1]["rows"], "fcfs_lanes": out[0]["rows"]}, fh)
    res = {k: v for k, v in out[1].items() if k != "rows"}
    res.update({"unit": "path samples/s", "requests": len(trace), "path_samples": total,
                "qps": len(trace) / trace[-1][0],
                "plain_fcfs": {"value": total / span_plain, "span_s": span_plain, "busy_s": sum(service),
                               "latency_ms": {"p50": _pct(lat_plain, 50) * 1e3, "p99": _pct(lat_plain, 99) * 1e3}},
                "express_off": {k: v for k, v in out[0].items() if k not in ("rows", "express_requests")},
                "service_s": service})
    return res


def serve_scenes(sphere):
    from hikari_amd import scenes
    return scenes.cbox(), sphere, scenes.heightfield()


def bench_batch(a):
    import hikari_amd as hk
    from hikari_amd import scenes
    ctx = hk.Context(0)
    gs = ctx.scene(scenes.displaced_sphere(alpha=a.alpha))
    gs.render(4)  # warm-up (kernels loaded, workspace allocated)
    b = batch_leg(gs, a.requests or 64, a.schedule, a.detail)
    out = {"metric": METRIC, "value": b["value"], "unit": "path samples/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": b["batched_s"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded procedural scene and request mix)",
           "config": {"workload": f"{b['requests']} batched requests over the displaced icosphere-8 scene, res "
                                  "{64..1024}^2, spp {16..1024}, camera on a radius-4 sphere"},
           "batch": b}
    print(json.dumps(out), flush=True)


def bench_serve(a):
    import hikari_amd as hk
    from hikari_amd import scenes, workloads
    box, sphere, field = serve_scenes(scenes.displaced_sphere(alpha=a.alpha))
    ctx = hk.Context(0)
    ctx.set_option("max_paths", a.max_paths or (1 << 26))  # plain renders at the server's bulk wavefront
    gss = [ctx.scene(box), ctx.scene(sphere), ctx.scene(field)]
    for g in gss:
        g.render(1)
    trace = workloads.serve_trace([box, None, field], n=a.requests or 40, qps=a.qps)
    s = serve_leg(ctx, gss, trace, a.budget_gb, a.detail)
    s.pop("service_s")
    out = {"metric": METRIC, "value": s["value"], "unit": "path samples/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": s["span_s"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded procedural scenes, Poisson request trace)",
           "config": {"workload": f"serving {len(trace)} requests at {a.qps} req/s over 3 resident scenes "
                                  "(box 5k tris, displaced sphere 1.3M, heightfield 2.1M); res {64..1024}^2, "
                                  "spp {16..1024}"},
           "serve": s}
    print(json.dumps(out), flush=True)


def bench_accel(a):
    """Batched rayIntersect (hk_accel_intersect_device) over the displaced-sphere
    mesh: 32 M rays leaving points just above the surface in uniform random
    directions, resident in HBM; closest-hit and any-hit rays per second."""
    import torch
    import hikari_amd as hk
    from hikari_amd import scenes
    scene = scenes.displaced_sphere(alpha=a.alpha)
    v = np.concatenate([m.v for m in scene.meshes]).astype(np.float32)
    f, base = [], 0
    for m in scene.meshes:
        f.append(m.f.astype(np.int64) + base)
        base += m.v.shape[0]
    f = np.concatenate(f).astype(np.uint32)
    ctx = hk.Context(0)
    acc = ctx.accel(v, f)
    n = 1 << 25
    g = torch.Generator(device="cuda").manual_seed(5)
    o = torch.nn.functional.normalize(torch.randn(n, 3, device="cuda", generator=g), dim=1) * 1.1
    d = torch.nn.functional.normalize(torch.randn(n, 3, device="cuda", generator=g), dim=1)
    rays = torch.cat([o, torch.full((n, 1), 1e-4, device="cuda"), d,
                      torch.full((n, 1), float("inf"), device="cuda")], dim=1).contiguous()
    hits = torch.empty((n, 4), device="cuda")
    out = {}
    for any_hit in (False, True):
        acc.intersect_device(rays.data_ptr(), hits.data_ptr(), n, any_hit)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            acc.intersect_device(rays.data_ptr(), hits.data_ptr(), n, any_hit)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        out["any_hit" if any_hit else "closest_hit"] = {"rays_per_s": n / dt, "ms": dt * 1e3,
                                                        "hit_rate": float((hits[:, 3].view(torch.int32) >= 0)
                                                                          .float().mean())}
    line = {
        "metric": "batched rayIntersect rays/s (hk_accel_intersect_device)", "value": out["closest_hit"]["rays_per_s"],
        "unit": "rays/s", "n_gpus": 1, "steps": a.steps, "warmup": 1, "ms_per_step": out["closest_hit"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded rays from a sphere of radius 1.1 in uniform directions)",
        "config": {"workload": f"{n} rays over the displaced icosphere-8 mesh ({f.shape[0]} tris), rays in HBM"},
        "closest_hit": out["closest_hit"], "any_hit": out["any_hit"],
    }
    print(json.dumps(line), flush=True)


def bench_copy(a):
    """Device-to-device copy of 256 MiB (the runtime's copy kernel): the
    reference launch for scaling the FETCH_SIZE / WRITE_SIZE counters of the
    render kernels to bytes, and the copy bandwidth the HBM roofline is read
    against."""
    import torch
    n = 256 << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    print(json.dumps({"metric": "device copy GB/s (read + write)", "value": 2 * n / dt / 1e9, "unit": "GB/s",
                      "n_gpus": 1, "steps": a.steps, "warmup": 1, "ms_per_step": dt * 1e3, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
                      "config": {"workload": "256 MiB device-to-device copy"}}), flush=True)


def relmse_leg(scene_fn, alpha):
    """The workload's scene on a reduced film (64 x 48, 16 spp): GPU vs the
    oracle at equal per-path streams (same random paths), so that the
    benchmark's kernels and scene data are checked for matched quality on
    every run."""
    import hikari_amd as hk
    from oracle import pyoracle as O
    sc = scene_fn(alpha, 64, 48)
    ctx = hk.Context(0)
    img, _ = ctx.scene(sc).render(16, seed=77)
    ctx.close()
    ref, _, _ = O.render(sc, req_seed=77, spp=(0, 16))
    lum = np.array([0.212671, 0.715160, 0.072169])
    rm = float(np.mean((img.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + 1e-2)))
    return {"film": "64x48", "spp": 16, "relmse_vs_oracle": rm,
            "mean_ratio": float((img @ lum).mean() / (ref @ lum).mean())}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU sharding is described in DESIGN.md and not run", file=sys.stderr)
        sys.exit(2)
    if a.workload == "batch":
        return bench_batch(a)
    if a.workload == "serve":
        return bench_serve(a)
    if a.workload == "accel":
        return bench_accel(a)
    if a.workload == "copy":
        return bench_copy(a)
    import hikari_amd as hk
    from hikari_amd import scenes, workloads

    t = time.time()
    if a.workload == "c2":
        scene = scenes.heightfield()
        if a.spp == 256:
            a.spp = 512
        workload = ("1024^2 heightfield (2,093,058 tris) in 8 patches of 5 materials (rough dielectric, rough "
                    f"plastic, coated rough conductor, dielectric, diffuse), 2 quad lights + env, 1280x720, {a.spp} spp")

        def scene_fn(alpha, w, h):
            return scenes.heightfield(width=w, height=h)
    else:
        scene = scenes.displaced_sphere(alpha=a.alpha)
        workload = ("displaced icosphere-8 (1,310,720 tris), GGX rough conductor alpha "
                    f"{a.alpha}, env map 512x256, 1024x768, {a.spp} spp, plain gpuPathMIS")

        def scene_fn(alpha, w, h):
            return scenes.displaced_sphere(alpha=alpha, width=w, height=h)
    log(f"scene built in {time.time() - t:.1f} s: {scene.n_triangles} triangles")
    ctx = hk.Context(0)
    if a.max_paths:
        ctx.set_option("max_paths", a.max_paths)
    ctx.set_option("bvh_builder", 1 if a.bvh_builder == "device" else 0)
    if a.bvh_cell_bits >= 0:
        ctx.set_option("bvh_cell_bits", a.bvh_cell_bits)
    ctx.set_option("graph", a.graph)
    ctx.set_option("refill", a.refill)
    ctx.set_option("refill_shadow", a.refill_shadow)
    if a.fetch_batch:
        ctx.set_option("fetch_batch", a.fetch_batch)
    t = time.time()
    gs = ctx.scene(scene)
    upload_s = time.time() - t
    log(f"scene resident in HBM ({gs.device_bytes() / 2**20:.0f} MiB) after {upload_s:.2f} s "
        f"({a.bvh_builder} BVH build)")
    tune = None
    if a.autotune and not a.max_paths:
        tune = gs.autotune(a.spp, seed=999)
        log(f"autotune: max_paths {tune['max_paths']}, fetch_batch {tune['fetch_batch']}, refill "
            f"{tune['refill']}/{tune['refill_shadow']}, tile {tune['tile']}, grid {tune['grid']} in "
            f"{tune['total_ms']:.0f} ms")
    for i in range(a.warmup):
        _, st = gs.render(a.spp, seed=1000 + i)
        log(f"warmup {i}: {st['gpu_ms']:.1f} ms")
    lat, tot = [], {"paths": 0, "extension_rays": 0, "shadow_rays": 0, "trace_ms": 0.0, "trace_launches": 0,
                    "gpu_ms": 0.0, "node_visits": 0, "triangle_tests": 0, "bounces_launched": 0, "graph_bounces": 0,
                    "invalid_samples": 0}
    t_all = time.perf_counter()
    for i in range(a.steps):
        t0 = time.perf_counter()
        img, st = gs.render(a.spp, seed=i)
        lat.append(time.perf_counter() - t0)
        for k in tot:
            tot[k] += st[k]
        log(f"step {i}: {lat[-1] * 1e3:.1f} ms, {st['paths'] / lat[-1]:.3e} paths/s")
    wall = time.perf_counter() - t_all
    value = tot["paths"] / wall
    rays = tot["extension_rays"] + tot["shadow_rays"]
    # The roofline of k_extend comes from one more render of the same request,
    # outside the timed region, with HIP events around every k_extend launch
    # on the context stream (profiled renders launch eagerly).
    ctx.set_option("profile", 1)
    _, pst = gs.render(a.spp, seed=a.steps)
    ctx.set_option("profile", 0)
    ext_launches = max(1, pst["trace_launches"])
    ext_avg_ms = pst["trace_ms"] / ext_launches
    rays_per_launch = pst["extension_rays"] / ext_launches
    achieved = EXTEND_B * rays_per_launch / (ext_avg_ms * 1e-3) / 1e9 if ext_avg_ms > 0 else 0.0
    render_bytes = SEG_B * tot["extension_rays"] + SHADOW_B * tot["shadow_rays"] + PATH_B * tot["paths"]
    roofline = {"bound": "hbm", "kernel": "k_extend", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "launch_ms": ext_avg_ms,
                "rays_per_launch": rays_per_launch,
                "timing": "HIP events around each k_extend launch of one extra (eager) render"}
    # Figures that need rocprofv3 counter passes (HBM bytes, BVH work per
    # ray, VALU issue) come from the committed summaries of those passes,
    # labelled with the commit they were measured on; traffic is scaled to
    # this run's rays per launch.
    fp = {}
    tp = os.path.join(ROOT, "profiles", "k_extend_traffic.json")
    if os.path.exists(tp):
        with open(tp) as fh:
            tj = json.load(fh)
        wl = "c2" if a.workload == "c2" else "c1"
        b_ray = tj.get("per_kernel", {}).get(f"{wl}_k_extend", {}).get("hbm_bytes_per_ray", tj.get("hbm_bytes_per_ray"))
        if b_ray:
            roofline["traffic"] = b_ray * rays_per_launch
            fp["traffic"] = {"hbm_bytes_per_ray": b_ray, "commit": tj.get("commit"),
                             "source": "profiles/k_extend_traffic.json (rocprofv3 FETCH_SIZE / WRITE_SIZE)"}
    tv = os.path.join(ROOT, "profiles", "k_extend_traversal.json")
    if os.path.exists(tv) and ext_avg_ms > 0:
        with open(tv) as fh:
            tj = json.load(fh)
        per = tj.get("c2" if a.workload == "c2" else "c1")
        if per:
            b_ray = (EXTEND_B + tj["node_bytes"] * per["node_visits_per_ray"]
                     + tj["triangle_bytes"] * per["triangle_tests_per_ray"])
            eff = b_ray * rays_per_launch / (ext_avg_ms * 1e-3) / 1e9
            fp["effective"] = {"bytes_per_ray": b_ray, "achieved": eff, "frac": eff / HBM_PEAK_GBS,
                               "commit": tj.get("commit"), "source": "profiles/k_extend_traversal.json"}
    vp = os.path.join(ROOT, "profiles", "valu_issue_c1_c2.json")
    if os.path.exists(vp):
        with open(vp) as fh:
            vj = json.load(fh)
        ke = vj.get("c2" if a.workload == "c2" else "c1", {}).get("k_extend")
        if ke:
            fp["issue"] = {"valu_issue_busy": ke["valu_issue_busy"], "lane_use": ke["lane_use"],
                           "commit": vj.get("commit"), "source": "profiles/valu_issue_c1_c2.json (rocprofv3 SQ)"}
    roofline["from_profiles"] = fp or None
    out = {
        "metric": METRIC, "value": value, "unit": "path samples/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scenes and request mixes)",
        "config": {"workload": workload, "spp": a.spp, "width": scene.width, "height": scene.height,
                   "triangles": scene.n_triangles, "bvh_builder": a.bvh_builder},
        "scene_create_s": upload_s,
        "autotune": None if tune is None else {k: tune[k] for k in ("max_paths", "fetch_batch", "refill",
                                                                    "refill_shadow", "grid", "tile", "total_ms")},
        "bounces": tot["bounces_launched"], "graph_bounces": tot["graph_bounces"],
        "invalid_samples": tot["invalid_samples"],
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
        "rays_per_s": rays / wall,
        "rays_per_path": rays / max(1, tot["paths"]),
        "roofline": roofline,
        "render_roofline_frac": render_bytes / wall / (HBM_PEAK_GBS * 1e9),
    }
    if tot["node_visits"]:  # builds with traversal counters only
        out["traversal"] = {"node_visits_per_ray": tot["node_visits"] / max(1, rays),
                            "triangle_tests_per_ray": tot["triangle_tests"] / max(1, rays)}
    t = time.time()
    out["quality"] = relmse_leg(scene_fn, a.alpha)
    log(f"relMSE vs oracle on the reduced film: {out['quality']['relmse_vs_oracle']:.2e} "
        f"({time.time() - t:.1f} s)")
    if a.workload == "c1" and not a.no_batch:
        out["batch"] = batch_leg(gs, a.requests or 64, a.schedule, a.detail)
    if a.workload == "c1" and not a.no_serve:
        # plain renders of the trace at the server's bulk wavefront (64 M
        # paths, its default; set here, it is also the server's)
        ctx.set_option("max_paths", 1 << 26)
        box, sphere, field = serve_scenes(scene)
        gss = [ctx.scene(box), gs, ctx.scene(field)]
        for g in gss:
            g.render(1)
        trace = workloads.serve_trace([box, None, field], n=40, qps=a.qps)
        out["serve"] = serve_leg(ctx, gss, trace, a.budget_gb)
        # the same requests arriving 3x faster: the GPU saturated
        trace = workloads.serve_trace([box, None, field], n=40, qps=3 * a.qps)
        out["serve_saturated"] = serve_leg(ctx, gss, trace, a.budget_gb, service=out["serve"]["service_s"])
        for k in ("serve", "serve_saturated"):
            out[k].pop("service_s")
        ctx.set_option("max_paths", 0)
        out["c0"] = c0_leg(gss[0])
        log(f"c0: {out['c0']['value']:.3e} path samples/s, {out['c0']['ms_per_render']:.1f} ms per render")
    if not a.no_cpu_baseline:
        t = time.time()
        out["cpu_baseline"] = cpu_baseline(scene, a.cpu_seconds)
        log(f"cpu baseline in {time.time() - t:.1f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
