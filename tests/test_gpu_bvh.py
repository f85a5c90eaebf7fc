"""GPU checks of the device BVH builder (hk_set_option "bvh_builder" = 1):
closest hits equal to brute force bit for bit, any-hit agreement, images
identical to those over the host-built BVH, and the tree's shape."""
import numpy as np
import pytest

from hikari_amd import HkError, scenes
from oracle import pyoracle as O
from tests.scenes_testlib import box_rays, grazing_rays, scene_triangles, scene_vf

pytestmark = pytest.mark.gpu


@pytest.fixture
def dctx(ctx):
    ctx.set_option("bvh_builder", 1)
    yield ctx
    ctx.set_option("bvh_builder", 0)


def _hits_equal(g, o):
    gt, gu, gv, gp = g
    ot, ou, ov, op = o
    hit = op >= 0
    same = gp.astype(np.int64) == op
    bits = lambda a: np.asarray(a, np.float32).view(np.uint32)  # noqa: E731
    return same & (~hit | ((bits(gt) == bits(ot)) & (bits(gu) == bits(ou)) & (bits(gv) == bits(ov))))


def _sphere(subdiv=6):
    v, f = scenes.icosphere(subdiv, 1.0)
    v = v * (1.0 + 0.08 * scenes.fbm3(v, 1))[:, None]
    f = scenes._ensure_outward(v, f, (0, 0, 0))
    return v.astype(np.float32), f.astype(np.uint32)


def test_device_bvh_sphere_bitexact(dctx):
    """~82k-triangle displaced icosphere: closest hits (t, u, v, prim) over the
    device-built BVH equal the oracle's brute-force-equivalent hits bit for bit."""
    rng = np.random.default_rng(31)
    v, f = _sphere()
    acc = dctx.accel(v, f)
    info = acc.info()
    print("device BVH:", info)
    tris = v[f].reshape(-1, 9).astype(np.float32)
    rays = np.concatenate([box_rays(rng, 100000, -1.5, 1.5), grazing_rays(rng, tris, 100000)])
    ok = _hits_equal(acc.intersect(rays), O.intersect(tris, rays, use_bvh=True))
    print("mismatches:", int((~ok).sum()), "of", len(ok))
    assert ok.all()
    assert info["n_triangles"] == len(f) and 0 < info["n_nodes"] < len(f)
    assert info["n_leaves"] >= len(f) / 8 and info["max_depth"] <= 31


@pytest.mark.parametrize("cell_bits", [0, 3, 18, 30])
def test_device_bvh_cell_bits_bitexact(dctx, cell_bits):
    """The SAH upper tree over Morton cells of any size (0: none, 3: a few
    cells, 18: cells of a triangle or two, 30: the largest setting) changes
    the tree, never the hits."""
    rng = np.random.default_rng(40 + cell_bits)
    v, f = _sphere(5)
    dctx.set_option("bvh_cell_bits", cell_bits)
    try:
        acc = dctx.accel(v, f)
    finally:
        dctx.set_option("bvh_cell_bits", 24)
    info = acc.info()
    print("cell bits", cell_bits, "device BVH:", info)
    tris = v[f].reshape(-1, 9).astype(np.float32)
    rays = np.concatenate([box_rays(rng, 40000, -1.5, 1.5), grazing_rays(rng, tris, 40000)])
    assert _hits_equal(acc.intersect(rays), O.intersect(tris, rays, use_bvh=True)).all()
    assert info["n_triangles"] == len(f) and info["max_depth"] <= 31
    with pytest.raises(HkError):
        dctx.set_option("bvh_cell_bits", 31)


def test_device_bvh_cbox_closest_and_any_hit(dctx):
    rng = np.random.default_rng(32)
    sc = scenes.cbox()
    v, f = scene_vf(sc)
    acc = dctx.accel(v, f)
    tris = scene_triangles(sc)
    rays = np.concatenate([box_rays(rng, 60000, -1.2, 1.2), grazing_rays(rng, tris, 40000)])
    assert _hits_equal(acc.intersect(rays), O.intersect(tris, rays, use_bvh=False)).all()
    rays[:, 7] = rng.uniform(0.05, 2.0, len(rays))
    g = acc.intersect(rays, any_hit=True)
    o = O.intersect(tris, rays, any_hit=True, use_bvh=False)
    assert np.array_equal(g[3] >= 0, o[3] >= 0)


@pytest.mark.parametrize("n_tris", [2, 3, 9, 17])
def test_device_bvh_tiny_meshes(dctx, n_tris):
    """Tiny inputs (a root whose children are leaves; a single-leaf tree)."""
    rng = np.random.default_rng(n_tris)
    tris = rng.uniform(-1, 1, (n_tris, 9)).astype(np.float32)
    v = tris.reshape(-1, 3)
    f = np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)
    acc = dctx.accel(v, f)
    rays = box_rays(rng, 20000, -1.2, 1.2)
    assert _hits_equal(acc.intersect(rays), O.intersect(tris, rays, use_bvh=False)).all()


def test_device_bvh_images_identical(ctx):
    """The heightfield (five materials, lights, environment) renders to the
    same bits over either builder's BVH: hits do not depend on the tree."""
    sc = scenes.heightfield(width=48, height=32, n=48)
    img_h, _ = ctx.scene(sc).render(8, seed=4)
    ctx.set_option("bvh_builder", 1)
    try:
        img_d, _ = ctx.scene(sc).render(8, seed=4)
    finally:
        ctx.set_option("bvh_builder", 0)
    assert np.array_equal(img_h.view(np.uint32), img_d.view(np.uint32))
