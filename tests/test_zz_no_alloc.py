"""hk_accel_intersect_device_noalloc against hk_accel_intersect_device.

The entry point that allocates nothing keeps the persistent intersect
kernel's fetch counter in the library's fixed per-device scratch instead of
the accelerator's own buffer.  Two calls in a row, the second on fewer rays,
must give the same closest hits bit for bit as the allocating entry point
(and the same occlusion answers for any-hit rays), write nothing past their
n hits, and a call beyond the ray limit must be refused on the host.

Device buffers come from the HIP runtime that libhikari_amd itself links
(hipMalloc through ctypes), not from torch: torch's own ROCm runtime has
failed to initialise in a process where the library had already run the
rest of the GPU suite.
"""
import ctypes as C
import os

import numpy as np
import pytest

import hikari_amd as hk
from hikari_amd import scenes
from tests.scenes_testlib import box_rays, grazing_rays

pytestmark = pytest.mark.gpu

SENTINEL = 0x7FC0DEAD  # a NaN bit pattern no hit record carries


def _sphere():
    v, f = scenes.icosphere(6, 1.0)
    v = v * (1.0 + 0.08 * scenes.fbm3(v, 1))[:, None]
    f = scenes._ensure_outward(v, f, (0, 0, 0))
    v = v.astype(np.float32)
    return v, f.astype(np.uint32), v[f].reshape(-1, 9)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library is linked to."""

    def __init__(self):
        hk.lib()
        self.L = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed with hipError {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self._ok(self.L.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.bufs.append(p)
        self._ok(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def download(self, ptr, like):
        out = np.empty_like(like)
        self._ok(self.L.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2), "hipMemcpy D2H")
        return out

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _run(ctx, acc, hip, rays_dev, n, cap, any_hit, noalloc):
    init = np.full((cap, 4), SENTINEL, np.int32)
    hits = hip.upload(init)
    call = acc.intersect_device_noalloc if noalloc else acc.intersect_device
    call(rays_dev, hits, n, any_hit=any_hit)
    ctx.sync()
    return hip.download(hits, init)


@pytest.mark.parametrize("any_hit", [False, True])
def test_noalloc_equals_allocating_entry_point(any_hit):
    rng = np.random.default_rng(71 + int(any_hit))
    v, f, tris = _sphere()
    ctx = hk.Context(0)
    hip = _Hip()
    try:
        acc = ctx.accel(v, f)
        n1, n2 = 150000, 20000
        rays = np.concatenate([box_rays(rng, n1 // 2, -1.5, 1.5), grazing_rays(rng, tris, n1 - n1 // 2)])
        rng.shuffle(rays)
        if any_hit:
            rays[:, 7] = rng.uniform(0.05, 2.0, n1)
        rays_dev = hip.upload(rays)
        for n in (n1, n2):
            print(f"noalloc any_hit={any_hit} n={n}", flush=True)
            ref = _run(ctx, acc, hip, rays_dev, n, n1, any_hit, noalloc=False)
            got = _run(ctx, acc, hip, rays_dev, n, n1, any_hit, noalloc=True)
            assert (got[n:] == SENTINEL).all(), "hits written past n"
            assert (got[:n, 3] != SENTINEL).all(), "rays left without a hit record"
            if any_hit:
                assert np.array_equal(got[:n, 3] >= 0, ref[:n, 3] >= 0)
            else:
                bad = ~(got[:n] == ref[:n]).all(axis=1)
                print("closest-hit mismatches:", int(bad.sum()), "of", n, "hit rate",
                      float((ref[:n, 3] >= 0).mean()))
                assert not bad.any()
        with pytest.raises(hk.HkError) as e:
            acc.intersect_device_noalloc(rays_dev, rays_dev, (1 << 31) + 1)
        assert e.value.code == -1 and "limit" in str(e.value)
    finally:
        hip.close()
        ctx.close()
