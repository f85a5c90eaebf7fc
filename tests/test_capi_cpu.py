"""CPU checks of the C ABI library: it loads and exports every entry point
include/hikari_amd.h declares (no compute calls here: there is no GPU)."""
import ctypes as C
import os
import re

import pytest

import hikari_amd as hk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "hikari_amd.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(hk_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "hk_accel_intersect" in syms and "hk_render" in syms and "hk_scene_create" in syms
    assert set(syms) == set(hk.EXPORTS)


@pytest.mark.skipif(not os.path.exists(hk.LIB_PATH), reason="libhikari_amd.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(hk.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(hk.LIB_PATH), reason="libhikari_amd.so not built")
def test_version_string():
    assert b"gfx950" in hk.lib().hk_version()


@pytest.mark.skipif(not os.path.exists(hk.LIB_PATH), reason="libhikari_amd.so not built")
def test_context_creation_fails_loudly_without_device():
    """No CPU fallback: without a usable HIP device the library refuses to create a context."""
    import subprocess
    import sys
    code = ("import hikari_amd as hk\n"
            "try:\n    hk.Context(0)\nexcept hk.HkError as e:\n    print('ERR', e.code)\nelse:\n    print('OK')\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, env=env,
                         timeout=120).stdout
    assert "ERR" in out


@pytest.mark.skipif(not os.path.exists(hk.LIB_PATH), reason="libhikari_amd.so not built")
def test_server_entry_points_refuse_null_handles():
    """The server entry points check their handle before touching a device."""
    L = hk.lib()
    assert L.hk_server_cancel(None, 1) == -1
    assert L.hk_server_wait(None, 1, None) == -1
    assert L.hk_server_get_stats(None, None) == -1
