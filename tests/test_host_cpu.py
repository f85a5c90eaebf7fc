"""CPU checks of the C++ host layer (hikari_amd/host): the XML scene loader
(checked against the independent test-side reader tests/xmlscene.py), its
error behaviour, and the EXR / PFM codecs (checked against numpy's float16
conversion and against the reference's own EXR renders).  No device is used:
the loader runs with --dump, images with --convert."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from hikari_amd import desc as D
from tests.xmlscene import parse_test

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
# HIKARI_RENDER selects another build of the driver (e.g. a sanitizer build)
CLI = os.environ.get("HIKARI_RENDER", os.path.join(ROOT, "hikari_amd", "hikari_render"))

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="hikari_render not built")


def run(*args, check=True):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    if check and r.returncode != 0:
        raise AssertionError(r.stderr)
    return r


def dump(path):
    return json.loads(run(path, "--dump").stdout)


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        a = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return a[::-1].copy()


def write_pfm(path, img):
    h, w, _ = img.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1\n".encode())
        f.write(np.ascontiguousarray(img[::-1], "<f4").tobytes())


@pytest.mark.parametrize("name", ["test-direct.xml", "test-furnace.xml"])
def test_loader_matches_test_reader(name):
    """Every scene of the reference's test file: same meshes (vertex / face
    counts, bounds), material records, emitters, camera and integrator as the
    Python reader builds."""
    path = os.path.join(GOLDEN, name)
    d = dump(path)
    refs, scenes = parse_test(path)
    assert np.allclose(d["references"], refs, rtol=1e-6)
    assert len(d["scenes"]) == len(scenes)
    for ds, ps in zip(d["scenes"], scenes):
        assert ps.integrator in ds["integrator"] and f"depth = {ps.max_depth}" in ds["integrator"]
        cam = ds["camera"]
        assert (cam["width"], cam["height"]) == (ps.width, ps.height)
        assert np.isclose(cam["fov"], ps.fov)
        assert np.allclose(np.array(cam["to_world"]).reshape(4, 4), ps.cam_to_world, atol=1e-5)
        assert len(ds["meshes"]) == len(ps.meshes)
        for dm, pm in zip(ds["meshes"], ps.meshes):
            assert dm["nv"] == pm.v.shape[0] and dm["nt"] == pm.f.shape[0]
            assert np.allclose(dm["lo"], pm.v.min(0), atol=1e-5) and np.allclose(dm["hi"], pm.v.max(0), atol=1e-5)
            rec = np.array(dm["materials"][-D.MAT_STRIDE:], np.float32)
            assert np.allclose(rec, ps.materials[pm.material], rtol=1e-6)
            if pm.radiance is not None:
                assert np.allclose(dm["radiance"], pm.radiance)
            else:
                assert "radiance" not in dm


def test_loader_reads_reference_scene_with_transforms():
    """odyssey-mis.xml (scene/pa5/odyssey): translate / scale chains applied in
    document order, each premultiplied (src/core/Parser.cpp:308-369)."""
    d = dump(os.path.join(GOLDEN, "odyssey", "odyssey-mis.xml"))
    assert d["spp"] == 32 and (d["camera"]["width"], d["camera"]["height"]) == (768, 432)
    # cube.obj spans [-0.5, 0.5]^3 (checked here) -> translate(0,0.5,0), scale(1,9,4)
    v, f, n, uv = D.load_obj(os.path.join(GOLDEN, "odyssey", "cube.obj"))
    assert np.allclose(v.min(0), -0.5) and np.allclose(v.max(0), 0.5)
    mono, floor, wall = d["meshes"]
    assert np.allclose(mono["lo"], [-0.5, 0, -2], atol=1e-5) and np.allclose(mono["hi"], [0.5, 9, 2], atol=1e-5)
    assert np.allclose(floor["lo"], [-12.5, -1, -9], atol=1e-5) and np.allclose(floor["hi"], [12.5, 0, 9], atol=1e-5)
    assert np.allclose(wall["lo"], [-12.5, 0, -9], atol=1e-5) and np.allclose(wall["hi"], [-11.5, 10, 9], atol=1e-5)
    assert wall["radiance"] == [1, 1, 1]


@pytest.mark.parametrize("xml,needle", [
    ("<scene><bogus/></scene>", "unexpected tag"),
    ('<scene><integrator type="pathMIS"><integer name="depth" value="x"/></integrator></scene>', "integer"),
    ('<scene><integrator type="pathMIS"/></scene>', "Property 'depth' is missing"),
    ('<scene><integrator type="nosuch"/></scene>', "could not be found"),
    ('<scene><integrator type="pathMIS"><integer name="depth" value="3" extra="1"/></integrator></scene>',
     "Unexpected attribute"),
    ('<scene><integrator type="pathMIS"><float name="depth" value="3"/></integrator></scene>', "wrong type"),
    ("<scene><translate value='1,2,3'/></scene>", "transform"),
    ('<scene><camera type="perspective"><color name="x" value="1, 2"/></camera></scene>', "Expected 3 values"),
    ("<scene><mesh type='obj'>", "missing closing tag"),
])
def test_loader_errors(tmp_path, xml, needle):
    p = tmp_path / "bad.xml"
    p.write_text(xml)
    r = run(str(p), "--dump", check=False)
    assert r.returncode != 0 and needle in r.stderr, r.stderr


_OPT_SCENE = """<scene>
  <integrator type="gpuPathMIS"><integer name="depth" value="100"/>{opts}</integrator>
  <camera type="perspective"><integer name="width" value="8"/><integer name="height" value="6"/></camera>
  <sampler type="independent"><integer name="sampleCount" value="4"/></sampler>
</scene>"""


def test_loader_reads_integrator_options(tmp_path):
    """queueSize and rrStartDepth, integrator properties the reference does
    not have, reach the GPU integrator (its description in --dump); both
    are optional and refuse negative values."""
    p = tmp_path / "opts.xml"
    p.write_text(_OPT_SCENE.format(opts=""))
    assert dump(str(p))["integrator"] == "GPUPathIntegrator[pathMIS, depth = 100]"
    p.write_text(_OPT_SCENE.format(opts='<integer name="rrStartDepth" value="3"/>'
                                        '<integer name="queueSize" value="1048576"/>'))
    assert dump(str(p))["integrator"] == "GPUPathIntegrator[pathMIS, depth = 100, rrStartDepth = 3, queueSize = 1048576]"
    for opt in ("rrStartDepth", "queueSize"):
        p.write_text(_OPT_SCENE.format(opts=f'<integer name="{opt}" value="-1"/>'))
        r = run(str(p), "--dump", check=False)
        assert r.returncode != 0 and f"{opt} must be >= 0" in r.stderr, r.stderr


def test_loader_requires_camera(tmp_path):
    p = tmp_path / "nocam.xml"
    p.write_text('<scene><integrator type="pathMIS"><integer name="depth" value="3"/></integrator></scene>')
    r = run(str(p), "--dump", check=False)
    assert r.returncode != 0 and "No camera" in r.stderr


def test_exr_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.uniform(-2, 40, (17, 23, 3)).astype(np.float32)
    img[0, 0] = [0.0, np.float32(1e-38), np.float32(3e38)]
    write_pfm(tmp_path / "a.pfm", img)
    run("--convert", str(tmp_path / "a.pfm"), str(tmp_path / "b.exr"))
    run("--convert", str(tmp_path / "b.exr"), str(tmp_path / "c.pfm"))
    back = read_pfm(tmp_path / "c.pfm")
    assert np.array_equal(back.view(np.uint32), img.view(np.uint32))


def _write_half_exr(path, img16):
    """Uncompressed scanline EXR with HALF R, G, B channels (OpenEXR layout)."""
    h, w, _ = img16.shape

    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data
    ch = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", 1, 0, 1, 1) for c in "BGR") + b"\0"
    hdr = struct.pack("<II", 20000630, 2)
    hdr += attr("channels", "chlist", ch) + attr("compression", "compression", b"\0")
    hdr += attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    hdr += attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, w - 1, h - 1))
    hdr += attr("lineOrder", "lineOrder", b"\0") + attr("pixelAspectRatio", "float", struct.pack("<f", 1))
    hdr += attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0)) + attr("screenWindowWidth", "float",
                                                                              struct.pack("<f", 1)) + b"\0"
    table_at = len(hdr)
    body, offsets = b"", []
    for y in range(h):
        offsets.append(table_at + 8 * h + len(body))
        row = b"".join(np.ascontiguousarray(img16[y, :, c], "<f2").tobytes() for c in (2, 1, 0))
        body += struct.pack("<ii", y, len(row)) + row
    with open(path, "wb") as f:
        f.write(hdr + struct.pack(f"<{h}Q", *offsets) + body)


def test_exr_half_channels_match_numpy_float16(tmp_path):
    rng = np.random.default_rng(6)
    vals = np.concatenate([rng.uniform(-60000, 60000, 300), rng.uniform(-1e-4, 1e-4, 300),  # normals / subnormals
                           [0.0, -0.0, 65504.0, 6e-8, np.inf, -np.inf]]).astype(np.float16)
    img16 = np.resize(vals, (11, 19 * 3)).reshape(11, 19, 3)
    _write_half_exr(tmp_path / "h.exr", img16)
    run("--convert", str(tmp_path / "h.exr"), str(tmp_path / "h.pfm"))
    back = read_pfm(tmp_path / "h.pfm")
    assert np.array_equal(back.view(np.uint32), img16.astype(np.float32).view(np.uint32))


def test_exr_decodes_reference_renders(tmp_path):
    """scene/pa5/odyssey/odyssey-{mis,ems,mats}.exr (ZIP-compressed FLOAT,
    written by OpenEXR): three independent estimators of one scene; decoded
    correctly they agree to within their Monte Carlo noise."""
    imgs = {}
    for k in ("mis", "ems", "mats"):
        run("--convert", os.path.join(GOLDEN, "odyssey", f"odyssey-{k}.exr"), str(tmp_path / f"{k}.pfm"))
        imgs[k] = read_pfm(tmp_path / f"{k}.pfm")
        assert imgs[k].shape == (432, 768, 3) and np.isfinite(imgs[k]).all() and imgs[k].min() >= 0
    lum = {k: (v @ np.array([0.212671, 0.715160, 0.072169], np.float32)) for k, v in imgs.items()}
    for k in ("ems", "mats"):
        assert abs(lum[k].mean() / lum["mis"].mean() - 1) < 0.02
        # 24x24-pixel block means agree far better than single pixels
        b = lambda a: a.reshape(18, 24, 32, 24).mean(axis=(1, 3))  # noqa: E731
        rel = np.abs(b(lum[k]) - b(lum["mis"])) / (b(lum["mis"]) + 0.05)
        assert np.median(rel) < 0.03, (k, float(np.median(rel)))


T_CASES = [  # mean, variance, reference, n, level, tests
    (1.0000000119, 0.0, 0.99999998, 10000, 0.01, 2),  # a deterministic estimator: variance floored
    (1.05, 0.5, 1.0, 1000, 0.01, 1),
    (1.1, 0.5, 1.0, 1000, 0.01, 1),
    (0.5, 2e-6, 0.4999, 100000, 0.01, 5),               # below the floor: t uses 1e-5
    (3.0, 4.0, 3.2, 50, 0.05, 15),
    (0.2, 1e-3, 0.2, 100000, 0.01, 6),
    (0.7, 0.09, 0.69, 100000, 0.01, 6),
]


@pytest.mark.parametrize("case", T_CASES)
def test_student_t_matches_scipy(case):
    """The harness's t-test (hikari_render --t-test) against scipy's Student t:
    t = |mean - reference| sqrt(n / max(variance, 1e-5)) on n - 1 degrees of
    freedom, two-sided p-value, Sidak level 1 - (1 - level)^(1 / tests),
    rejected iff p < level (ext/hypothesis/hypothesis.h:319-351)."""
    from scipy import stats
    mean, var, ref, n, level, tests = case
    r = run("--t-test", *(repr(x) for x in case), check=False)
    t = abs(mean - ref) * np.sqrt(n / max(var, 1e-5))
    p = 2 * stats.t.sf(t, n - 1)
    alpha = 1 - (1 - level) ** (1 / tests)
    got = dict(kv.split(" = ") for kv in r.stdout.split("(")[1].rstrip(")\n").replace("level ", "level = ").split(", "))
    assert float(got["t"]) == pytest.approx(t, rel=1e-5)
    assert float(got["p-value"]) == pytest.approx(p, rel=1e-4, abs=1e-12)
    assert float(got["level"]) == pytest.approx(alpha, rel=1e-5)
    assert r.returncode == (1 if p < alpha else 0)


@pytest.mark.parametrize("x,dof", [(0.5, 1), (3.5, 4), (180.0, 199), (250.0, 199), (1e-3, 2), (40.0, 10)])
def test_chi2_cdf_matches_scipy(x, dof):
    """The chi-square distribution function of the chi2test harness
    (hikari_render --chi2-cdf) against scipy's."""
    from scipy import stats
    got = float(run("--chi2-cdf", repr(x), str(dof)).stdout)
    assert got == pytest.approx(stats.chi2.cdf(x, dof), rel=1e-10, abs=1e-14)


def _chi2_table(tmp_path, obs, exp, count, min_exp=5, level=0.01, tests=1):
    p = tmp_path / "table.txt"
    p.write_text(f"{len(obs)} {count} {min_exp} {level} {tests}\n" + " ".join(map(repr, map(float, obs))) + "\n"
                 + " ".join(map(repr, map(float, exp))) + "\n")
    r = run("--chi2-table", str(p), check=False)
    assert r.returncode in (0, 1), r.stderr
    return r


def _chi2_fields(out):
    import re
    m = re.search(r"chi\^2 = ([^,]+), (\d+) d\.o\.f\. \((\d+) cells pooled\), p-value = ([^,]+),", out)
    return float(m.group(1)), int(m.group(2)), int(m.group(3)), float(m.group(4))


@pytest.mark.parametrize("seed", range(4))
def test_pooled_chi2_test_matches_scipy(tmp_path, seed):
    """The pooled chi-square test of the chi2test harness (hikari_render
    --chi2-table, ext/hypothesis/hypothesis.h:140-240): without low cells it
    is Pearson's test on k - 1 degrees of freedom (scipy.stats.chisquare);
    with cells below the minimum expected frequency, those are merged, in
    increasing expected frequency, until the merged cell reaches it."""
    from scipy import stats
    rng = np.random.default_rng(seed)
    n = 5000
    # no pooling: 40 cells, all expected frequencies >= 5
    p = rng.uniform(0.5, 2.0, 40)
    p /= p.sum()
    obs = rng.multinomial(n, p)
    r = _chi2_table(tmp_path, obs, p * n, n)
    chsq, dof, pooled, pv = _chi2_fields(r.stdout)
    ref = stats.chisquare(obs, p * n)
    assert pooled == 0 and dof == 39
    assert chsq == pytest.approx(ref.statistic, rel=1e-9) and pv == pytest.approx(ref.pvalue, rel=1e-6, abs=1e-12)
    # pooling: 12 cells of tiny expected frequency among 30
    q = np.concatenate([rng.uniform(1e-5, 4e-4, 12), rng.uniform(0.5, 2.0, 18)])
    q /= q.sum()
    obs = rng.multinomial(n, q)
    e = q * n
    order = np.argsort(e, kind="stable")
    k = 0
    pool_e = 0.0
    while k < len(order) and (e[order[k]] < 5 or 0 < pool_e < 5):
        pool_e += e[order[k]]
        k += 1
    keep = order[k:]
    merged_o = np.concatenate([obs[keep], [obs[order[:k]].sum()]])
    merged_e = np.concatenate([e[keep], [e[order[:k]].sum()]])
    ref = stats.chisquare(merged_o, merged_e)
    r = _chi2_table(tmp_path, obs, e, n)
    chsq, dof, pooled, pv = _chi2_fields(r.stdout)
    assert pooled == k and dof == len(merged_o) - 1
    assert chsq == pytest.approx(ref.statistic, rel=1e-9) and pv == pytest.approx(ref.pvalue, rel=1e-6, abs=1e-12)


def test_pooled_chi2_test_rejections(tmp_path):
    """Samples in a cell of expected frequency 0 beyond 1e-5 of the sample
    count reject (rounding may leave fewer); a clearly different
    distribution rejects; too few degrees of freedom reject."""
    n = 200000
    exp = [50000.0, 50000.0, 50000.0, 50000.0, 0.0]
    assert _chi2_table(tmp_path, [49990, 50010, 50000, 49997, 3], exp, n).returncode == 1
    assert _chi2_table(tmp_path, [49990, 50010, 50000, 49998, 2], exp, n).returncode == 0
    assert _chi2_table(tmp_path, [52000, 48000, 50000, 50000, 0], exp, n).returncode == 1
    r = _chi2_table(tmp_path, [10], [10.0], 10)
    assert r.returncode == 1 and "degrees of freedom" in r.stdout


def test_chi2test_file_loads():
    """scene/pa5/tests/chi2test-microfacet.xml: the chi2test type with the
    reference's defaults (resolution 10 x 20, 5 incident directions per BSDF,
    5000 samples per cell)."""
    d = json.loads(run(os.path.join(GOLDEN, "chi2test-microfacet.xml"), "--dump").stdout)
    assert d["test"] == "chi2test"
    assert "resolution = 10 x 20" in d["description"] and "sampleCount = 1000000" in d["description"]
    assert "testCount = 5" in d["description"] and "minExpFrequency = 5" in d["description"]
