"""hikari_amd: MI355X-native pathMIS core (HIP kernels for gfx950 behind a C ABI).

Python access goes through ctypes to libhikari_amd.so (include/hikari_amd.h).
There is no CPU fallback: importing works without a GPU, but creating a
Context fails loudly when the HIP library or device is missing.
"""
import ctypes as C
import os

import numpy as np

from . import desc  # noqa: F401
from .desc import SceneDesc

HERE = os.path.dirname(os.path.abspath(__file__))
# HIKARI_AMD_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("HIKARI_AMD_LIB", os.path.join(HERE, "libhikari_amd.so"))

HK_OK = 0
ERRORS = {-1: "HK_ERR_INVALID", -2: "HK_ERR_HIP", -3: "HK_ERR_OOM", -4: "HK_ERR_UNSUPPORTED", -5: "HK_ERR_IO",
          -6: "HK_ERR_BUDGET", -7: "HK_ERR_CANCELLED"}
HK_ERR_CANCELLED = -7

EXPORTS = ["hk_ctx_create", "hk_ctx_destroy", "hk_last_error", "hk_version", "hk_ctx_sync", "hk_accel_build",
           "hk_accel_destroy", "hk_accel_get_info", "hk_accel_intersect", "hk_accel_intersect_device",
           "hk_accel_intersect_device_noalloc",
           "hk_bsdf_batch", "hk_scene_create", "hk_scene_destroy", "hk_scene_device_bytes", "hk_render",
           "hk_render_pixel_samples", "hk_render_batch", "hk_set_option", "hk_server_create", "hk_server_add_scene",
           "hk_server_submit", "hk_server_wait", "hk_server_cancel", "hk_server_get_stats", "hk_server_destroy", "hk_autotune"]


class HkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Ray(C.Structure):
    _fields_ = [("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("tmin", C.c_float),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float), ("tmax", C.c_float)]


class AccelInfo(C.Structure):
    _fields_ = [("n_triangles", C.c_uint32), ("n_nodes", C.c_uint32), ("n_leaves", C.c_uint32),
                ("max_depth", C.c_uint32), ("node_bytes", C.c_uint64), ("triangle_bytes", C.c_uint64),
                ("build_ms", C.c_double)]


class Request(C.Structure):
    _fields_ = [("cam_to_world", C.c_float * 16), ("fov", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("spp", C.c_uint32), ("spp_offset", C.c_uint32), ("seed", C.c_uint32),
                ("use_scene_camera", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("bounces_launched", C.c_uint64), ("gpu_ms", C.c_double), ("trace_ms", C.c_double),
                ("trace_launches", C.c_uint64), ("node_visits", C.c_uint64), ("triangle_tests", C.c_uint64),
                ("graph_bounces", C.c_uint64), ("invalid_samples", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("queue_ms", C.c_double), ("latency_ms", C.c_double), ("paths", C.c_uint64),
                ("submit_ms", C.c_double), ("lane", C.c_int32), ("invalid_samples", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("bounces", C.c_uint64), ("paths", C.c_uint64),
                ("fixed_bytes", C.c_uint64), ("film_bytes", C.c_uint64), ("shared_bounces", C.c_uint64),
                ("max_in_flight", C.c_uint64), ("cancelled", C.c_uint64), ("requests_in_flight", C.c_uint64)]


_lib = None


def lib():
    """Load libhikari_amd.so (built by __graft_entry__.build() / make -C hikari_amd/csrc)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C hikari_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.hk_last_error.restype = C.c_char_p
        L.hk_last_error.argtypes = [P]
        L.hk_version.restype = C.c_char_p
        L.hk_ctx_create.argtypes = [C.c_int, C.POINTER(P)]
        L.hk_ctx_destroy.argtypes = [P]
        L.hk_ctx_sync.argtypes = [P]
        L.hk_set_option.argtypes = [P, C.c_char_p, C.c_int64]
        L.hk_accel_build.argtypes = [P, P, C.c_uint32, P, C.c_uint32, C.POINTER(P)]
        L.hk_accel_destroy.argtypes = [P]
        L.hk_accel_get_info.argtypes = [P, C.POINTER(AccelInfo)]
        L.hk_accel_intersect.argtypes = [P, P, C.c_uint32, P, C.c_int]
        L.hk_accel_intersect_device.argtypes = [P, P, C.c_uint32, P, C.c_int]
        L.hk_accel_intersect_device_noalloc.argtypes = [P, P, C.c_uint32, P, C.c_int]
        L.hk_bsdf_batch.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint32, P, P, P, P, P, P, P, P, P,
                                    P, P]
        L.hk_scene_create.argtypes = [P, C.POINTER(SceneDesc), C.POINTER(P)]
        L.hk_scene_destroy.argtypes = [P]
        L.hk_scene_device_bytes.argtypes = [P]
        L.hk_scene_device_bytes.restype = C.c_uint64
        L.hk_render.argtypes = [P, C.POINTER(Request), P, C.POINTER(RenderStats)]
        L.hk_render_pixel_samples.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P]
        L.hk_render_batch.argtypes = [P, C.POINTER(Request), C.c_uint32, C.POINTER(C.POINTER(C.c_float)),
                                      C.c_uint32, C.POINTER(RenderStats)]
        L.hk_server_create.argtypes = [P, C.c_uint64, C.POINTER(P)]
        L.hk_server_add_scene.argtypes = [P, P, C.POINTER(C.c_int32)]
        L.hk_server_submit.argtypes = [P, C.c_int32, C.POINTER(Request), P, C.POINTER(C.c_uint64)]
        L.hk_server_wait.argtypes = [P, C.c_uint64, C.POINTER(ServerResult)]
        L.hk_server_cancel.argtypes = [P, C.c_uint64]
        L.hk_server_get_stats.argtypes = [P, C.POINTER(ServerStats)]
        L.hk_server_destroy.argtypes = [P]
        L.hk_autotune.argtypes = [P, C.POINTER(Request), C.c_uint64, C.POINTER(TuneResult)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One device + one HIP stream (hk_ctx)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().hk_ctx_create(device, C.byref(self.h))
        if rc != HK_OK:
            raise HkError(rc, lib().hk_last_error(None).decode())

    def check(self, rc):
        if rc != HK_OK:
            raise HkError(rc, lib().hk_last_error(self.h).decode())
        return rc

    def set_option(self, name, value):
        self.check(lib().hk_set_option(self.h, name.encode(), int(value)))

    def sync(self):
        """Wait for the context stream (hk_ctx_sync)."""
        self.check(lib().hk_ctx_sync(self.h))

    def close(self):
        if self.h:
            lib().hk_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- acceleration
    def accel(self, v, f):
        return Accel(self, v, f)

    # ---- BSDF batch
    def bsdf_batch(self, materials, index, wi, wo, measure, u, rft_dir=None):
        mats = np.ascontiguousarray(materials, np.float32)
        wi = np.ascontiguousarray(wi, np.float32)
        wo = np.ascontiguousarray(wo, np.float32)
        measure = np.ascontiguousarray(measure, np.int32)
        u = np.ascontiguousarray(u, np.float32)
        n = wi.shape[0]
        out = dict(eval=np.zeros((n, 3), np.float32), pdf=np.zeros(n, np.float32), f=np.zeros((n, 3), np.float32),
                   wo=np.zeros((n, 3), np.float32), measure=np.zeros(n, np.int32), eta=np.zeros(n, np.float32),
                   spdf=np.zeros(n, np.float32))
        self.check(lib().hk_bsdf_batch(self.h, _p(mats), mats.shape[0], index,
                                       rft_dir.encode() if rft_dir else None, n, _p(wi), _p(wo), _p(measure), _p(u),
                                       _p(out["eval"]), _p(out["pdf"]), _p(out["f"]), _p(out["wo"]),
                                       _p(out["measure"]), _p(out["eta"]), _p(out["spdf"])))
        return out

    def scene(self, scene):
        return GpuScene(self, scene)


class Accel:
    """Batched rayIntersect over a 4-wide BVH (hk_accel)."""

    def __init__(self, ctx, v, f):
        self.ctx = ctx
        v = np.ascontiguousarray(v, np.float32)
        f = np.ascontiguousarray(f, np.uint32)
        self.h = C.c_void_p()
        ctx.check(lib().hk_accel_build(ctx.h, _p(v), v.shape[0], _p(f), f.shape[0], C.byref(self.h)))

    def info(self):
        i = AccelInfo()
        self.ctx.check(lib().hk_accel_get_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in AccelInfo._fields_}

    def intersect(self, rays, any_hit=False):
        """rays: (n, 8) float32 (o, tmin, d, tmax) -> (t, u, v, prim) arrays."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        self.ctx.check(lib().hk_accel_intersect(self.h, _p(rays), n, _p(hits), int(any_hit)))
        prim = hits[:, 3].view(np.int32).copy()
        return hits[:, 0].copy(), hits[:, 1].copy(), hits[:, 2].copy(), prim

    def intersect_device(self, rays_ptr, hits_ptr, n, any_hit=False):
        """Device arrays (e.g. torch tensors' data_ptr()): rays n x 8 float32,
        hits n x 4; runs on the context's stream without waiting for it."""
        self.ctx.check(lib().hk_accel_intersect_device(self.h, C.c_void_p(rays_ptr), int(n), C.c_void_p(hits_ptr),
                                                       int(any_hit)))

    def intersect_device_noalloc(self, rays_ptr, hits_ptr, n, any_hit=False):
        """intersect_device through the entry point that allocates nothing (its
        fetch counter in the library's fixed per-device scratch)."""
        self.ctx.check(lib().hk_accel_intersect_device_noalloc(self.h, C.c_void_p(rays_ptr), int(n),
                                                               C.c_void_p(hits_ptr), int(any_hit)))

    def __del__(self):
        try:
            if self.h:
                lib().hk_accel_destroy(self.h)
        except Exception:
            pass


def make_request(scene=None, spp=1, seed=0, spp_offset=0, cam_to_world=None, fov=None, width=None, height=None):
    r = Request()
    r.spp, r.seed, r.spp_offset = spp, seed, spp_offset
    if cam_to_world is None:
        r.use_scene_camera = 1
        if scene is not None:
            r.width, r.height, r.fov = scene.width, scene.height, scene.fov
            r.cam_to_world[:] = [float(x) for x in np.asarray(scene.cam_to_world, np.float64).ravel()]
    else:
        r.use_scene_camera = 0
        r.cam_to_world[:] = [float(x) for x in np.asarray(cam_to_world, np.float64).ravel()]
        r.fov, r.width, r.height = fov, width, height
    return r


TUNE_MAX = 32


class TuneResult(C.Structure):
    _fields_ = [("max_paths", C.c_uint32), ("grid", C.c_uint32), ("tile", C.c_uint32), ("refill", C.c_uint32),
                ("refill_shadow", C.c_uint32), ("fetch_batch", C.c_uint32), ("n_tried", C.c_uint32),
                ("tried_max_paths", C.c_uint32 * TUNE_MAX), ("tried_grid", C.c_uint32 * TUNE_MAX),
                ("tried_tile", C.c_uint32 * TUNE_MAX), ("tried_refill", C.c_uint32 * TUNE_MAX),
                ("tried_refill_shadow", C.c_uint32 * TUNE_MAX), ("tried_fetch_batch", C.c_uint32 * TUNE_MAX),
                ("samples_per_s", C.c_double * TUNE_MAX), ("total_ms", C.c_double)]


class GpuScene:
    """A scene resident in HBM (hk_scene)."""

    def __init__(self, ctx, scene):
        self.ctx = ctx
        self.scene = scene
        d = scene.to_c()
        self._desc = d
        self.h = C.c_void_p()
        ctx.check(lib().hk_scene_create(ctx.h, C.byref(d), C.byref(self.h)))

    def device_bytes(self):
        return lib().hk_scene_device_bytes(self.h)

    def render(self, spp, seed=0, spp_offset=0, request=None):
        req = request or make_request(self.scene, spp, seed, spp_offset)
        W, H = (self.scene.width, self.scene.height) if req.use_scene_camera else (req.width, req.height)
        img = np.zeros((H, W, 3), np.float32)
        st = RenderStats()
        self.ctx.check(lib().hk_render(self.h, C.byref(req), _p(img), C.byref(st)))
        return img, st.as_dict()

    def render_batch(self, requests, max_paths=0):
        n = len(requests)
        arr = (Request * n)(*requests)
        imgs = []
        ptrs = (C.POINTER(C.c_float) * n)()
        for i, r in enumerate(requests):
            W, H = (self.scene.width, self.scene.height) if r.use_scene_camera else (r.width, r.height)
            im = np.zeros((H, W, 3), np.float32)
            imgs.append(im)
            ptrs[i] = im.ctypes.data_as(C.POINTER(C.c_float))
        st = RenderStats()
        self.ctx.check(lib().hk_render_batch(self.h, arr, n, ptrs, max_paths, C.byref(st)))
        return imgs, st.as_dict()

    def autotune(self, spp, seed=0, calib_samples=0, request=None):
        """hk_autotune: picks the context's max_paths / grid / tile for this
        kind of request (results never depend on them) and returns the choice
        with every candidate's calibration throughput."""
        req = request or make_request(self.scene, spp, seed)
        t = TuneResult()
        self.ctx.check(lib().hk_autotune(self.h, C.byref(req), C.c_uint64(calib_samples), C.byref(t)))
        tried = [{"max_paths": t.tried_max_paths[i], "grid": t.tried_grid[i], "tile": t.tried_tile[i],
                  "refill": t.tried_refill[i], "refill_shadow": t.tried_refill_shadow[i],
                  "fetch_batch": t.tried_fetch_batch[i], "samples_per_s": t.samples_per_s[i]}
                 for i in range(t.n_tried)]
        return {"max_paths": t.max_paths, "grid": t.grid, "tile": t.tile, "refill": t.refill,
                "refill_shadow": t.refill_shadow, "fetch_batch": t.fetch_batch, "total_ms": t.total_ms,
                "tried": tried}

    def pixel_samples(self, x, y, s0, n, seed=0):
        out = np.zeros((n, 5), np.float32)
        self.ctx.check(lib().hk_render_pixel_samples(self.h, seed, x, y, s0, n, _p(out)))
        return out

    def __del__(self):
        try:
            if self.h:
                lib().hk_scene_destroy(self.h)
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over resident scenes (hk_server)."""

    def __init__(self, ctx, hbm_budget=0):
        self.ctx = ctx
        self.h = C.c_void_p()
        self._scenes = []
        self._pending = {}
        ctx.check(lib().hk_server_create(ctx.h, int(hbm_budget), C.byref(self.h)))

    def add_scene(self, gpu_scene):
        sid = C.c_int32()
        self.ctx.check(lib().hk_server_add_scene(self.h, gpu_scene.h, C.byref(sid)))
        self._scenes.append(gpu_scene)
        return sid.value

    def submit(self, scene_id, request):
        gs = self._scenes[scene_id]
        W, H = (gs.scene.width, gs.scene.height) if request.use_scene_camera else (request.width, request.height)
        img = np.zeros((H, W, 3), np.float32)
        t = C.c_uint64()
        self.ctx.check(lib().hk_server_submit(self.h, scene_id, C.byref(request), _p(img), C.byref(t)))
        self._pending[t.value] = img
        return t.value

    def wait(self, ticket):
        """-> (image, {"status", "queue_ms", "latency_ms", "paths"})"""
        res = ServerResult()
        rc = lib().hk_server_wait(self.h, ticket, C.byref(res))
        img = self._pending.pop(ticket, None)
        self.ctx.check(rc)
        return img, {k: getattr(res, k) for k, _ in ServerResult._fields_}

    def cancel(self, ticket):
        """hk_server_cancel: the ticket's wait then raises HK_ERR_CANCELLED
        (unless the request had already finished)."""
        self.ctx.check(lib().hk_server_cancel(self.h, ticket))

    def stats(self):
        st = ServerStats()
        self.ctx.check(lib().hk_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().hk_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
