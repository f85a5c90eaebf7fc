"""CPU check of the emitter-CDF guide tables (hikari_amd/csrc/hk_cdf.h): the
guided search used by the kernels returns, for every CDF shape and sample
tried (random samples, every CDF value and its neighbouring floats, every
bucket boundary, zero / negative / NaN / infinite / >= 1 samples), the index
of DiscretePDF1D's search over the whole CDF (src/core/DiscretePDF.cpp:32-78).
A table with its entries moved by one must fail the same check."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("cdf") / "cdf_guide_check"
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cdf_guide_check.cpp"), "-o", str(exe)],
                   check=True)
    return str(exe)


def _run(exe, *args):
    out = subprocess.run([exe, *args], capture_output=True, text=True, check=True, timeout=300).stdout.split()
    return int(out[1]), int(out[3])


def test_guided_search_equals_full_search(checker):
    checked, bad = _run(checker)
    assert checked > 1_000_000 and bad == 0


def test_shifted_guide_is_caught(checker):
    checked, bad = _run(checker, "--plant")
    assert bad > 0
