"""Writes tests/golden/veach_mi/: the scene files of the reference's
scene/pa5/veach_mi (XML + OBJ, copied as data) and veach-mats-blocks16.npy,
the 16 x 16-pixel block means (RGB, float32, 32 x 48 blocks) of the
reference's own render veach-mats.exr, decoded with this project's EXR reader
(hikari_render --convert).  Run once in a tree where the reference is
present: python tests/golden/make_veach_fixture.py [REFERENCE_ROOT]."""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CLI = os.path.join(ROOT, "hikari_amd", "hikari_render")
OUT = os.path.join(HERE, "veach_mi")


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        a = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return a[::-1].copy()


def main(ref_root):
    src = os.path.join(ref_root, "scene", "pa5", "veach_mi")
    os.makedirs(OUT, exist_ok=True)
    for name in sorted(os.listdir(src)):
        if name.endswith((".xml", ".obj")):
            shutil.copyfile(os.path.join(src, name), os.path.join(OUT, name))
    with tempfile.TemporaryDirectory() as tmp:
        pfm = os.path.join(tmp, "ref.pfm")
        subprocess.run([CLI, "--convert", os.path.join(src, "veach-mats.exr"), pfm], check=True)
        img = read_pfm(pfm)
    h, w, _ = img.shape
    blocks = img.reshape(h // 16, 16, w // 16, 16, 3).mean(axis=(1, 3)).astype(np.float32)
    np.save(os.path.join(OUT, "veach-mats-blocks16.npy"), blocks)
    print("image", img.shape, "blocks", blocks.shape, "mean", img.mean(axis=(0, 1)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
