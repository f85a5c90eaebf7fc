"""GPU parity on the benchmark configurations at their real geometry.

* C1: the displaced icosphere-8 (1,310,722 triangles with the ground) under
  the 512 x 256 sun-patch environment map, GGX rough gold at each alpha of the
  workload, rendered on a reduced film against the oracle at equal streams;
  closest hits on the full mesh bit for bit.
* C2: the full 1024^2 heightfield (2,093,058 triangles): closest hits bit for
  bit, and an image against the oracle at equal streams.
* C3: a slice of the 64-request mix over the C1 scene (random poses on the
  radius-4 sphere, mixed films and spp, scaled down) batched vs alone.
* C4: the server over the three resident scenes (box, displaced sphere,
  heightfield) with requests arriving while others run: each result equals
  the request alone, and the requests demonstrably shared the device.
"""
import copy
import time

import numpy as np
import pytest

import hikari_amd as hk
from hikari_amd import scenes, workloads
from oracle import pyoracle as O
from tests.scenes_testlib import box_rays, grazing_rays, lum, rel_mse, scene_triangles, scene_vf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_scene():
    return scenes.displaced_sphere(alpha=0.3, width=64, height=48)


@pytest.fixture(scope="module")
def c2_scene():
    return scenes.heightfield(width=64, height=36)


def _hits_equal(g, o):
    gt, gu, gv, gp = g
    ot, ou, ov, op = o
    same = gp.astype(np.int64) == op
    hit = op >= 0
    bits = ((gt.view(np.uint32) == ot.view(np.uint32)) & (gu.view(np.uint32) == ou.view(np.uint32))
            & (gv.view(np.uint32) == ov.view(np.uint32)))
    return same & (~hit | bits)


def _full_mesh_hits(ctx, scene, lo, hi, seed):
    rng = np.random.default_rng(seed)
    v, f = scene_vf(scene)
    tris = scene_triangles(scene)
    acc = ctx.accel(v, f)
    rays = np.concatenate([box_rays(rng, 12000, lo, hi), grazing_rays(rng, tris, 8000)])
    g = acc.intersect(rays)
    o = O.intersect(tris, rays, use_bvh=True)
    ok = _hits_equal(g, o)
    print(f"{len(tris)} triangles: closest-hit mismatches {int((~ok).sum())} of {len(ok)}, "
          f"hit rate {float((o[3] >= 0).mean()):.3f}", flush=True)
    assert ok.all()
    # a brute-force subset pins the oracle's own BVH on this mesh
    sub = rays[::100]
    b = O.intersect(tris, sub, use_bvh=False)
    assert _hits_equal(tuple(x[::100] for x in g), b).all()
    # any hit agrees on occlusion
    short = rays.copy()
    short[:, 7] = rng.uniform(0.01, 1.0, len(rays))
    ga = acc.intersect(short, any_hit=True)
    oa = O.intersect(tris, short, any_hit=True, use_bvh=True)
    assert np.array_equal(ga[3] >= 0, oa[3] >= 0)


def test_c1_full_mesh_hits_bitexact(ctx, c1_scene):
    """Closest hits over the full C1 geometry equal the oracle BVH (itself
    checked against brute force on a subset) in t, u, v and prim, bit for bit."""
    _full_mesh_hits(ctx, c1_scene, -1.3, 1.3, 71)


def test_c2_full_heightfield_hits_bitexact(ctx, c2_scene):
    """The same on the full C2 heightfield (2,093,058 triangles + lights)."""
    _full_mesh_hits(ctx, c2_scene, -6.0, 6.0, 72)


@pytest.mark.parametrize("alpha", [0.1, 0.3, 0.5])
def test_c1_image_vs_oracle(ctx, c1_scene, alpha):
    """C1 at each alpha of the workload (GGX rough gold, RGB eta / k, diffuse
    ground, the 512 x 256 sun-patch environment map sampled by its CDF):
    64 x 48 film, 16 spp, GPU vs the FP32 oracle at equal per-path streams:
    relMSE < 1e-3 and mean luminance within 1%."""
    sc = copy.deepcopy(c1_scene)
    metal = sc.meshes[0].material
    sc.materials[metal][22] = alpha
    sc.materials[metal][23] = alpha
    img, st = ctx.scene(sc).render(16, seed=31)
    ref, _, ost = O.render(sc, req_seed=31, spp=(0, 16))
    rm = rel_mse(img, ref)
    print(f"C1 alpha {alpha}: relMSE {rm:.3e}, mean {lum(img).mean():.5f} vs {lum(ref).mean():.5f}, "
          f"rays/path {(st['extension_rays'] + st['shadow_rays']) / st['paths']:.2f}", flush=True)
    assert np.isfinite(img).all()
    assert rm < 1e-3
    assert abs(lum(img).mean() - lum(ref).mean()) / lum(ref).mean() < 0.01


def test_c2_full_heightfield_image_vs_oracle(ctx, c2_scene):
    """C2 on the full heightfield (5 material kinds, 2 quad lights + env):
    64 x 36 film, 8 spp, GPU vs the FP32 oracle at equal streams."""
    img, st = ctx.scene(c2_scene).render(8, seed=32)
    ref, _, _ = O.render(c2_scene, req_seed=32, spp=(0, 8))
    rm = rel_mse(img, ref)
    print(f"C2 full heightfield: relMSE {rm:.3e}, mean {lum(img).mean():.5f} vs {lum(ref).mean():.5f}", flush=True)
    assert rm < 1e-3
    assert abs(lum(img).mean() - lum(ref).mean()) / lum(ref).mean() < 0.01


def _c3_slice(n=12):
    """The C3 mix's generator (random poses on the radius-4 sphere) with
    films {32..128}^2 and spp {1..8}: the same request kinds, scaled down."""
    return workloads.batch_requests(n=n, seed=2024, resolutions=(32, 48, 64, 128), spps=(1, 2, 4, 8))


def test_c3_batch_slice_equals_alone(ctx, c1_scene):
    """A slice of the C3 mix over the full C1 scene: the batch returns each
    request's film bit for bit as rendered alone, with the default wavefront
    and with one small enough to recycle path slots across requests."""
    gs = ctx.scene(c1_scene)
    reqs = _c3_slice()
    alone = [gs.render(0, request=r)[0] for r in reqs]
    for max_paths in (0, 20000):
        batch, st = gs.render_batch(reqs, max_paths=max_paths)
        assert st["paths"] == sum(r.spp * r.width * r.height for r in reqs)
        for a, b in zip(alone, batch):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_c3_request_vs_oracle(ctx, c1_scene):
    """One posed request of the C3 slice (its own camera on the radius-4
    sphere) against the oracle rendering the same camera at equal streams."""
    r = _c3_slice()[3]
    img, _ = ctx.scene(c1_scene).render(0, request=r)
    sc = copy.deepcopy(c1_scene)
    sc.cam_to_world = np.array(r.cam_to_world[:], np.float64).reshape(4, 4)
    sc.fov, sc.width, sc.height = r.fov, r.width, r.height
    ref, _, _ = O.render(sc, req_seed=r.seed, spp=(0, r.spp))
    rm = rel_mse(img, ref)
    print(f"C3 request {r.width}x{r.height}x{r.spp}: relMSE {rm:.3e}", flush=True)
    assert rm < 1e-3


def test_c4_server_three_scenes_equal_alone(c1_scene, c2_scene):
    """C4 over the three resident scenes (box, displaced sphere, heightfield):
    requests of the serving mix (scaled down) arrive while earlier ones run;
    each film equals the same request rendered alone, bit for bit, and the
    server's own clock shows requests admitted while others were in flight
    (their bounces carried more than one request)."""
    box = scenes.cbox(width=48, height=48)
    c = hk.Context(0)
    c.set_option("max_paths", 60000)  # several bounces per request, so that arrivals overlap
    gss = [c.scene(box), c.scene(c1_scene), c.scene(c2_scene)]
    trace = workloads.serve_trace([box, None, c2_scene], n=12, qps=200.0, seed=78, resolutions=(32, 48, 64),
                                  spps=(2, 4, 8, 16))
    alone = [gss[sid].render(0, request=r)[0] for _, sid, r in trace]
    srv = hk.Server(c)
    for g in gss:
        srv.add_scene(g)
    t0 = time.perf_counter()
    tickets = []
    for arr, sid, r in trace:
        d = t0 + arr - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets.append(srv.submit(sid, r))
    res = []
    for t, a, (_, sid, r) in zip(tickets, alone, trace):
        img, rr = srv.wait(t)
        assert rr["status"] == 0 and rr["paths"] == r.spp * r.width * r.height
        assert np.array_equal(img.view(np.uint32), a.view(np.uint32)), (sid, r.width, r.spp, r.seed)
        res.append(rr)
    st = srv.stats()
    srv.close()
    c.close()
    admit = [x["submit_ms"] + x["queue_ms"] for x in res]
    done = [x["submit_ms"] + x["latency_ms"] for x in res]
    overlaps = sum(1 for j in range(len(res)) for i in range(j) if admit[i] <= admit[j] < done[i])
    print(f"C4 slice: {len(res)} requests, {overlaps} admitted while an earlier one was in flight, "
          f"shared bounces {st['shared_bounces']} of {st['bounces']}, max in flight {st['max_in_flight']}", flush=True)
    assert st["completed"] == len(trace)
    assert overlaps > 0 and st["shared_bounces"] > 0 and st["max_in_flight"] >= 2
