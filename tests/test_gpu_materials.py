"""GPU checks of the mirror / microfacet / roughDiffuse / twoSided BSDFs and
the whitted estimator: per-path agreement with the oracle, closed forms, and
the reference's own render of scene/pa5/veach_mi (microfacet plates)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.scenes_testlib import env_plane_scene, lum, materials_box, rel_mse
from tests.xmlscene import parse_test

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("integrator", ["pathMIS", "pathEMS", "pathMATS", "whitted"])
def test_materials_box_vs_oracle(ctx, integrator):
    """Whole image at equal per-path streams, every estimator, on the box
    holding all four BSDFs."""
    sc = materials_box(integrator=integrator)
    gs = ctx.scene(sc)
    img, st = gs.render(16, seed=5)
    ref, _, _ = O.render(sc, req_seed=5, spp=(0, 16))
    rm = rel_mse(img, ref)
    print(integrator, "materials box relMSE vs oracle:", rm, st)
    assert rm < 1e-3
    assert abs(lum(img).mean() - lum(ref).mean()) / lum(ref).mean() < 0.01


def test_whitted_paths_track_oracle(ctx):
    """test-direct.xml scenes under whitted: the same (pixel, sample) keys
    give the same estimates on the GPU and in the oracle."""
    refs, scs = parse_test(os.path.join(GOLDEN, "test-direct.xml"))
    for sc in [s for s in scs if s.integrator == "pathMIS"]:
        sc.integrator, sc.max_depth = "whitted", -1
        g = ctx.scene(sc).pixel_samples(0, 0, 0, 3000)
        o = O.pixel_samples(sc, 0, 0, 0, 3000)
        assert np.array_equal(g[:, 3:], o[:, 3:])
        close = np.all(np.abs(g[:, :3] - o[:, :3]) <= 1e-4 * np.maximum(1.0, np.abs(o[:, :3])), axis=1)
        print("whitted identical estimates:", close.mean())
        assert close.mean() > 0.99


def _whitted(sc):
    sc.integrator, sc.max_depth = "whitted", -1
    return sc


def test_whitted_closed_forms_gpu(ctx):
    """Diffuse plane under a constant sky: a L (one environment sample per
    vertex); the same plane facing away: 0, or a L inside twoSided; a mirror:
    L / 0.95 with probability 0.95, else 0."""
    from hikari_amd import desc as D
    n = 200000
    v = lum(ctx.scene(_whitted(env_plane_scene(albedo=0.5))).pixel_samples(0, 0, 0, n)[:, :3])
    assert abs(v.mean() - 0.5) < 4 * v.std(ddof=1) / np.sqrt(n), v.mean()
    for two_sided, expect in ((False, 0.0), (True, 0.5)):
        sc = _whitted(env_plane_scene(albedo=0.5))
        sc.meshes[0].f = np.ascontiguousarray(sc.meshes[0].f[:, ::-1])
        if two_sided:
            sc.materials[sc.meshes[0].material] = D.two_sided(sc.add_material(D.diffuse(0.5)))
        v = lum(ctx.scene(sc).pixel_samples(0, 0, 0, n)[:, :3])
        assert abs(v.mean() - expect) < 4 * v.std(ddof=1) / np.sqrt(n) + 1e-9, (two_sided, v.mean())
    sc = _whitted(env_plane_scene(radiance=2.0))
    sc.materials[sc.meshes[0].material] = D.mirror()
    sc.cam_to_world = D.lookat((0.0, -3.0, 4.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    v = lum(ctx.scene(sc).pixel_samples(0, 0, 0, n)[:, :3])
    assert np.all(np.isclose(v, 2.0 / 0.95, rtol=1e-6) | (v == 0))
    assert abs((v == 0).mean() - 0.05) < 4 * np.sqrt(0.05 * 0.95 / n)
