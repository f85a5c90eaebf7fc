"""FP32 scene ingestion of the host layer against the reference's Eigen.

The reference composes a toWorld transform op by op on an Eigen::Affine3f
(src/core/Parser.cpp:308-369) and applies it to OBJ vertices and normals in
FP32 (src/mesh/WavefrontObjMesh.cpp:52-70, src/core/Transform.cpp:36-50).
oracle/_ref/eigen_xform does exactly that with the reference's vendored
Eigen 3.3.5; hikari_render --dump-mesh writes what the host layer built.
World-space positions and composed matrices must agree bit for bit; normals
go through the inverse, which the host computes in FP64 and rounds (the
reference uses Eigen's FP32 4x4 routine), so they agree to a few ulp.
"""
import json
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "oracle", "_ref", "eigen_xform")
# HIKARI_RENDER selects another build of the driver (e.g. a sanitizer build)
RENDER = os.environ.get("HIKARI_RENDER", os.path.join(ROOT, "hikari_amd", "hikari_render"))
GOLDEN = os.path.join(ROOT, "tests", "golden")

pytestmark = [pytest.mark.skipif(not os.path.exists(PROBE), reason="reference Eigen probe not built"),
              pytest.mark.skipif(not os.path.exists(RENDER), reason="hikari_render not built")]


def _tokens(s):
    return [t for t in s.replace(",", " ").split() if t]


def _ops(transform_el):
    """The transform element's ops in document order, as probe lines."""
    lines = []
    for op in transform_el:
        if op.tag in ("translate", "scale"):
            lines.append(op.tag + " " + " ".join(_tokens(op.get("value"))))
        elif op.tag == "rotate":
            lines.append("rotate " + op.get("angle").strip() + " " + " ".join(_tokens(op.get("axis"))))
        elif op.tag == "matrix":
            lines.append("matrix " + " ".join(_tokens(op.get("value"))))
        elif op.tag == "lookat":
            lines.append("lookat " + " ".join(_tokens(op.get("origin")) + _tokens(op.get("target"))
                                              + _tokens(op.get("up"))))
    return lines


def _probe(tmp_path, ops, obj):
    opf = tmp_path / "ops.txt"
    opf.write_text("\n".join(ops) + "\n")
    out = tmp_path / "probe.bin"
    subprocess.run([PROBE, str(opf), obj or "-", str(out)], check=True)
    raw = np.fromfile(out, np.uint8)
    f = raw.view(np.float32)
    m, inv = f[:16].reshape(4, 4), f[16:32].reshape(4, 4)
    nv = int(raw[128:132].view(np.int32)[0])
    pts = f[33:33 + 3 * nv].reshape(nv, 3)
    vecs = f[33 + 3 * nv:33 + 6 * nv].reshape(nv, 3)
    k = 33 + 6 * nv
    nn = int(f[k:k + 1].view(np.int32)[0])
    nrm = f[k + 1:k + 1 + 3 * nn].reshape(nn, 3)
    return m, inv, pts, vecs, nrm


def _dump(tmp_path, scene):
    prefix = str(tmp_path / "mesh")
    out = subprocess.run([RENDER, scene, "--dump-mesh", prefix], check=True, capture_output=True, text=True,
                         cwd=tmp_path).stdout
    desc = json.loads(out, parse_int=float)  # "-0" keeps its sign
    meshes = []
    for i in range(len(desc["meshes"])):
        v = np.fromfile(f"{prefix}.{i}.v.f32", np.float32).reshape(-1, 3)
        n = np.fromfile(f"{prefix}.{i}.n.f32", np.float32).reshape(-1, 3)
        meshes.append((v, n))
    return desc, meshes


def _rows(a):
    return {tuple(r) for r in np.ascontiguousarray(a, np.float32).view(np.uint32).reshape(-1, 3).tolist()}


def _ulps(a, b):
    ia = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


@pytest.mark.parametrize("scene", ["odyssey/odyssey-ems.xml", "veach_mi/veach-mis.xml"])
def test_reference_scene_vertices_bitexact(tmp_path, scene):
    """Every world-space vertex of the reference's own scenes equals the
    Eigen FP32 result bit for bit, and the camera's toWorld matrix too."""
    path = os.path.join(GOLDEN, scene)
    desc, meshes = _dump(tmp_path, path)
    root = ET.parse(path).getroot()
    objs = [m for m in root.iter("mesh") if m.get("type") == "obj"]
    assert len(objs) == len(meshes)
    checked = 0
    for el, (v, n) in zip(objs, meshes):
        fname = [s.get("value") for s in el.findall("string") if s.get("name") == "filename"][0]
        tr = [t for t in el.findall("transform") if t.get("name") == "toWorld"]
        _, _, pts, _, nrm = _probe(tmp_path, _ops(tr[0]) if tr else [], os.path.join(os.path.dirname(path), fname))
        host, ref = _rows(v), _rows(pts)
        assert host <= ref, f"{fname}: {len(host - ref)} of {len(host)} vertices differ from Eigen's"
        if len(n) and len(nrm):
            d = np.min(np.max(np.abs(n[:, None, :] - nrm[None, :, :]), axis=2), axis=1)
            assert d.max() <= 1e-6, (fname, float(d.max()))
        checked += len(host)
    assert checked > 0
    cam = root.find("camera")
    tr = [t for t in cam.findall("transform") if t.get("name") == "toWorld"]
    if tr:
        m, _, _, _, _ = _probe(tmp_path, _ops(tr[0]), None)
        got = np.array(desc["camera"]["to_world"], np.float32).reshape(4, 4)
        assert np.array_equal(got.view(np.uint32), m.view(np.uint32))


SCENE = """<scene>
  <integrator type="pathMIS"><integer name="depth" value="4"/></integrator>
  <camera type="perspective">
    <float name="fov" value="30"/>
    <integer name="width" value="8"/>
    <integer name="height" value="8"/>
    <transform name="toWorld">{cam}</transform>
  </camera>
  <mesh type="obj">
    <string name="filename" value="rand.obj"/>
    <transform name="toWorld">{ops}</transform>
    <bsdf type="diffuse"/>
    <emitter type="area"><color name="radiance" value="1, 1, 1"/></emitter>
  </mesh>
</scene>
"""


def _f(x):
    return repr(float(np.float32(x)))


def _random_ops(rng, k):
    xml, lines = [], []
    for _ in range(k):
        kind = rng.integers(0, 5)
        if kind == 0:
            v = [_f(x) for x in rng.normal(size=3) * 10]
            xml.append(f'<translate value="{", ".join(v)}"/>')
            lines.append("translate " + " ".join(v))
        elif kind == 1:
            v = [_f(x) for x in rng.uniform(0.1, 5, 3) * rng.choice([-1, 1], 3)]
            xml.append(f'<scale value="{", ".join(v)}"/>')
            lines.append("scale " + " ".join(v))
        elif kind == 2:
            a, ax = _f(rng.uniform(-180, 180)), [_f(x) for x in rng.normal(size=3)]
            xml.append(f'<rotate angle="{a}" axis="{", ".join(ax)}"/>')
            lines.append(f"rotate {a} " + " ".join(ax))
        elif kind == 3:
            m = rng.normal(size=(4, 4))
            m[3] = (0, 0, 0, 1)
            v = [_f(x) for x in m.ravel()]
            xml.append(f'<matrix value="{" ".join(v)}"/>')
            lines.append("matrix " + " ".join(v))
        else:
            o, t, u = ([_f(x) for x in rng.normal(size=3) * 5] for _ in range(3))
            xml.append(f'<lookat origin="{", ".join(o)}" target="{", ".join(t)}" up="{", ".join(u)}"/>')
            lines.append("lookat " + " ".join(o + t + u))
    return "".join(xml), lines


@pytest.mark.parametrize("seed", range(6))
def test_random_transforms_bitexact(tmp_path, seed):
    """Random chains of translate / scale / rotate (unnormalised axes, as the
    reference's AngleAxis takes them) / matrix / lookat: matrix, points and
    camera toWorld bit for bit, normals within 64 ulp or 1e-6."""
    rng = np.random.default_rng(100 + seed)
    n = 64
    v = rng.normal(size=(n, 3)) * 3
    vn = rng.normal(size=(n, 3))
    with open(tmp_path / "rand.obj", "w") as fo:
        for p in v:
            fo.write("v %s %s %s\n" % tuple(_f(x) for x in p))
        for p in vn:
            fo.write("vn %s %s %s\n" % tuple(_f(x) for x in p))
        for i in range(0, n - 2, 3):
            fo.write(f"f {i + 1}//{i + 1} {i + 2}//{i + 2} {i + 3}//{i + 3}\n")
    ops_xml, ops = _random_ops(rng, int(rng.integers(1, 6)))
    cam_xml, cam_ops = _random_ops(rng, 2)
    (tmp_path / "s.xml").write_text(SCENE.format(ops=ops_xml, cam=cam_xml))
    desc, meshes = _dump(tmp_path, str(tmp_path / "s.xml"))
    hv, hn = meshes[0]
    _, _, pts, _, nrm = _probe(tmp_path, ops, str(tmp_path / "rand.obj"))
    used = 3 * ((n - 2 + 2) // 3)
    assert hv.shape == (used, 3)
    assert np.array_equal(hv.view(np.uint32), pts[:used].view(np.uint32))
    u = _ulps(hn, nrm[:used])
    close = (u <= 64) | (np.abs(hn - nrm[:used]) <= 1e-6)
    assert close.all(), int(u.max())
    m, _, _, _, _ = _probe(tmp_path, cam_ops, None)
    got = np.array(desc["camera"]["to_world"], np.float32).reshape(4, 4)
    assert np.array_equal(got.view(np.uint32), m.view(np.uint32))


@pytest.mark.parametrize("seed", range(4))
def test_oracle_transform_products_match_eigen(tmp_path, seed):
    """The oracle's camera / environment transform products (M4::point, M4::vec)
    round like Eigen's Matrix4f * Vector4f and 3x3-block * Vector3f."""
    from oracle import pyoracle as O
    rng = np.random.default_rng(200 + seed)
    _, ops = _random_ops(rng, 3)
    v = rng.normal(size=(200, 3)) * 4
    with open(tmp_path / "p.obj", "w") as fo:
        for p in v:
            fo.write("v %s %s %s\n" % tuple(_f(x) for x in p))
    m, _, pts, vecs, _ = _probe(tmp_path, ops, str(tmp_path / "p.obj"))
    src = v.astype(np.float32)
    op, ov = O.xform(m, src)
    assert np.array_equal(op.view(np.uint32), pts.view(np.uint32))
    assert np.array_equal(ov.view(np.uint32), vecs.view(np.uint32))
