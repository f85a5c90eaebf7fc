"""Writes the reference-render fixtures of tests/golden/table and
tests/golden/envmap from the reference's scene files:

  * table/: scene/pa5/table (pathMIS, pathEMS and pathMATS over a microfacet
    table, a diffuse floor, water and glass dielectrics and two area lights):
    the three XML files and meshes/ copied as data, and table-<kind>-blocks20.npy,
    the 20 x 20-pixel block means (RGB, float32, 30 x 40 blocks) of the
    reference's own renders table-<kind>.exr (512 spp);
  * envmap/: scene/pa6 envmap0 and envmap3 (whitted, a diffuse sphere and
    plane under an environment map): the XML, OBJ and environment EXR files
    copied as data, and envmap<n>-blocks16.npy, the 16 x 16-pixel block means
    (32 x 32 blocks) of the reference's renders envmap<n>.exr (256 spp);
  * pa4tests/: scene/pa4/tests test-mesh.xml and test-mesh-furnace.xml
    (whitted t-tests with their reference values) and their meshes/;
  * chi2test-microfacet.xml: scene/pa5/tests' chi-square test of the
    microfacet BSDF's sampling (three parameter sets);
  * cbox4/ and motto/: scene/pa4 cbox-distributed and motto-diffuse /
    motto-dielectric (whitted over area lights, diffuse and dielectric):
    XML and meshes/, and <name>-blocks<b>.npy, the b x b-pixel block means of
    the reference's renders <name>.exr (512 spp; b = 20 for the 800 x 600
    cbox, 16 for the 768 x 384 motto).

Images are decoded with this project's EXR reader (hikari_render --convert).
Run once in a tree where the reference is present:
python tests/golden/make_reference_fixtures.py [REFERENCE_ROOT]."""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CLI = os.path.join(ROOT, "hikari_amd", "hikari_render")


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        a = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return a[::-1].copy()


def block_means(exr, b):
    with tempfile.TemporaryDirectory() as tmp:
        pfm = os.path.join(tmp, "ref.pfm")
        subprocess.run([CLI, "--convert", exr, pfm], check=True, capture_output=True)
        img = read_pfm(pfm)
    h, w, _ = img.shape
    return img.reshape(h // b, b, w // b, b, 3).mean(axis=(1, 3)).astype(np.float32)


def copy(src, dst, names):
    os.makedirs(dst, exist_ok=True)
    for name in names:
        shutil.copyfile(os.path.join(src, name), os.path.join(dst, name))


def main(ref_root):
    src = os.path.join(ref_root, "scene", "pa5", "table")
    out = os.path.join(HERE, "table")
    copy(src, out, [f"table-{k}.xml" for k in ("mis", "ems", "mats")])
    meshes = os.path.join(src, "meshes")
    copy(meshes, os.path.join(out, "meshes"), sorted(n for n in os.listdir(meshes) if n.endswith(".obj")))
    for k in ("mis", "ems", "mats"):
        blk = block_means(os.path.join(src, f"table-{k}.exr"), 20)
        np.save(os.path.join(out, f"table-{k}-blocks20.npy"), blk)
        print("table", k, blk.shape, blk.mean(axis=(0, 1)))
    src = os.path.join(ref_root, "scene", "pa6")
    out = os.path.join(HERE, "envmap")
    for n in (0, 3):
        copy(src, out, [f"envmap{n}.xml", f"envmap_{n}.exr"])
        blk = block_means(os.path.join(src, f"envmap{n}.exr"), 16)
        np.save(os.path.join(out, f"envmap{n}-blocks16.npy"), blk)
        print("envmap", n, blk.shape, blk.mean(axis=(0, 1)))
    copy(src, out, ["sphere.obj", "plane.obj"])
    copy(os.path.join(ref_root, "scene", "pa5", "tests"), HERE, ["chi2test-microfacet.xml"])
    src = os.path.join(ref_root, "scene", "pa4")
    copy(os.path.join(src, "tests"), os.path.join(HERE, "pa4tests"), ["test-mesh.xml", "test-mesh-furnace.xml"])
    for sub, out_name, names, b in (("cbox", "cbox4", ["cbox-distributed"], 20),
                                    ("motto", "motto", ["motto-diffuse", "motto-dielectric"], 16)):
        out = os.path.join(HERE, out_name)
        copy(os.path.join(src, sub), out, [n + ".xml" for n in names])
        for n in names:
            blk = block_means(os.path.join(src, sub, n + ".exr"), b)
            np.save(os.path.join(out, f"{n}-blocks{b}.npy"), blk)
            print(n, blk.shape, blk.mean(axis=(0, 1)))
    for sub, out_name in (("tests", "pa4tests"), ("cbox", "cbox4"), ("motto", "motto")):
        meshes = os.path.join(src, sub, "meshes")
        copy(meshes, os.path.join(HERE, out_name, "meshes"), sorted(n for n in os.listdir(meshes) if n.endswith(".obj")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
