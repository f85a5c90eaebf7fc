"""Request mixes of the batched and served workloads (seeded, synthetic).

Batched: 64 requests over one resident scene, camera on a sphere of radius 4
around the origin (elevation 5..80 degrees so that the ground plane faces
the camera), resolution drawn from {64, 128, 256, 512, 1024}^2 and spp from
{16, 64, 256, 1024}.  Served: Poisson arrivals at a fixed rate over several
resident scenes with the same per-request mix.
"""
import math

import numpy as np

from . import desc as D
from . import make_request

RESOLUTIONS = (64, 128, 256, 512, 1024)
SPPS = (16, 64, 256, 1024)


def random_pose(rng, radius=4.0, target=(0.0, -0.1, 0.0)):
    el = math.radians(rng.uniform(5.0, 80.0))
    az = rng.uniform(0.0, 2.0 * math.pi)
    o = (radius * math.cos(el) * math.cos(az), radius * math.sin(el), radius * math.cos(el) * math.sin(az))
    return D.lookat(o, target, (0.0, 1.0, 0.0))


def batch_requests(n=64, seed=2024, resolutions=RESOLUTIONS, spps=SPPS, fov=40.0):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        res = int(rng.choice(resolutions))
        spp = int(rng.choice(spps))
        reqs.append(make_request(cam_to_world=random_pose(rng), fov=fov, width=res, height=res, spp=spp,
                                 seed=1000 + i))
    return reqs


def serve_trace(scenes, n=40, qps=1.0, seed=77, resolutions=RESOLUTIONS, spps=SPPS):
    """-> list of (arrival_s, scene_id, request), arrivals Poisson at `qps`.

    `scenes` holds, per resident scene, either None (random pose on the radius-4
    sphere) or the scene itself (its own camera pose and field of view)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / qps)
        sid = int(rng.integers(0, len(scenes)))
        res = int(rng.choice(resolutions))
        spp = int(rng.choice(spps))
        sc = scenes[sid]
        pose = random_pose(rng) if sc is None else np.asarray(sc.cam_to_world, np.float64)
        fov = 40.0 if sc is None else float(sc.fov)
        out.append((t, sid, make_request(cam_to_world=pose, fov=fov, width=res, height=res, spp=spp, seed=5000 + i)))
    return out
