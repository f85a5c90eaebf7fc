"""ctypes wrapper of the CPU restatement (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; the product (hikari_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.ho_last_error.restype = C.c_char_p
        L.ho_pcg32_demo.argtypes = [C.c_int, C.c_char_p, C.c_int]
        L.ho_path_floats.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P]
        L.ho_vec_ops.argtypes = [C.c_int, P, P, P]
        L.ho_xform.argtypes = [P, C.c_int, P, P]
        L.ho_intersect.argtypes = [P, C.c_uint32, P, C.c_uint32, C.c_int, C.c_int, P, P, P, P]
        L.ho_bsdf_batch.argtypes = [C.c_int, P, C.c_int, C.c_int, C.c_char_p, C.c_int, P, P, P, P, P, P, P, P, P, P, P]
        L.ho_rft_reduce.argtypes = [C.c_char_p, C.c_float, C.c_float, C.c_int, P, P]
        L.ho_rft_eval.argtypes = [C.c_char_p, C.c_int, C.c_float, C.c_float, C.c_float]
        L.ho_rft_eval.restype = C.c_double
        L.ho_fresnel_dielectric.argtypes = [C.c_double, C.c_double, P]
        L.ho_fresnel_dielectric.restype = C.c_double
        L.ho_fresnel_conductor.argtypes = [C.c_double, P, P, P]
        L.ho_microfacet.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, P, P, P, P, P, P, P]
        L.ho_render.argtypes = [C.c_int, C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_int, P, P, P]
        L.ho_pixel_samples.argtypes = [C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.c_uint32, P]
        L.ho_rft_eval_diffuse.argtypes = [C.c_char_p, C.c_int, C.c_float, C.c_float]
        L.ho_rft_eval_diffuse.restype = C.c_double
        L.ho_set_slip.argtypes = [C.c_int]
        L.ho_film_splat.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                    P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc < 0:
        raise RuntimeError(lib().ho_last_error().decode())
    return rc


def pcg32_demo(rounds=5):
    buf = C.create_string_buffer(1 << 16)
    _check(lib().ho_pcg32_demo(rounds, buf, len(buf)))
    return buf.value.decode()


def path_floats(req_seed, pixel, sample, n):
    out = np.zeros(n, np.float32)
    lib().ho_path_floats(req_seed, pixel, sample, n, _p(out))
    return out


def xform(m, pts):
    """M4::point and M4::vec of the row-major 4x4 FP32 matrix m on pts (n x 3)."""
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    pts = np.ascontiguousarray(pts, np.float32)
    out = np.zeros((pts.shape[0], 6), np.float32)
    lib().ho_xform(_p(m), pts.shape[0], _p(pts), _p(out))
    return out[:, :3], out[:, 3:]


def vec_ops(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.zeros((a.shape[0], 13), np.float32)
    lib().ho_vec_ops(a.shape[0], _p(a), _p(b), _p(out))
    return out


def intersect(tris, rays, any_hit=False, use_bvh=True):
    """tris: (nt, 9) float32 (p0,p1,p2); rays: (n, 8) float32 (o, tmin, d, tmax)."""
    tris = np.ascontiguousarray(tris, np.float32)
    rays = np.ascontiguousarray(rays, np.float32)
    n = rays.shape[0]
    t, u, v = (np.zeros(n, np.float32) for _ in range(3))
    prim = np.zeros(n, np.int64)
    _check(lib().ho_intersect(_p(tris), tris.shape[0], _p(rays), n, int(any_hit), int(use_bvh), _p(t), _p(u), _p(v),
                              _p(prim)))
    return t, u, v, prim


def bsdf_batch(materials, index, wi, wo, measure, u, fp64=True, rft_dir=None):
    """Evaluate material `index` of the (m, 32) record table on n queries.

    Returns dict: eval (n,3), pdf (n), and Sample() with uniforms u[:, :2] (plus
    u[:, 2:4] for draws made inside Sample): f (n,3), wo (n,3), measure, eta, pdf.
    """
    from hikari_amd.desc import DATA_DIR
    mats = np.ascontiguousarray(materials, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    wo = np.ascontiguousarray(wo, np.float32)
    measure = np.ascontiguousarray(measure, np.int32)
    u = np.ascontiguousarray(u, np.float32)
    n = wi.shape[0]
    out = dict(eval=np.zeros((n, 3)), pdf=np.zeros(n), f=np.zeros((n, 3)), wo=np.zeros((n, 3)),
               measure=np.zeros(n, np.int32), eta=np.zeros(n), spdf=np.zeros(n))
    _check(lib().ho_bsdf_batch(int(fp64), _p(mats), mats.shape[0], index, (rft_dir or DATA_DIR).encode(), n, _p(wi),
                               _p(wo), _p(measure), _p(u), _p(out["eval"]), _p(out["pdf"]), _p(out["f"]),
                               _p(out["wo"]), _p(out["measure"]), _p(out["eta"]), _p(out["spdf"])))
    return out


def rft_reduce(path, eta, alpha, stage):
    t = np.zeros(50 * 50 * 100 * 2, np.float32)
    d = np.zeros(50 * 50 * 2, np.float32)
    n = _check(lib().ho_rft_reduce(path.encode(), eta, alpha, stage, _p(t), _p(d)))
    return t[:n], d


def rft_eval(path, stage, eta, alpha, cos_t):
    return lib().ho_rft_eval(path.encode(), stage, eta, alpha, cos_t)


def rft_eval_diffuse(path, stage, eta, alpha):
    return lib().ho_rft_eval_diffuse(path.encode(), stage, eta, alpha)


# Planted mistakes of the restatement (oracle/hk_oracle.hpp, enum Slip), for
# the tests that show a pinning check fails when the oracle is wrong.
SLIPS = ["none", "ggx_g1", "beckmann_g1", "conductor_fresnel", "rough_conductor_g", "rough_dielectric_eta",
         "plastic_fresnel", "plastic_eta2", "rough_plastic_eta", "coating_sigma", "coating_exit", "rough_coating_t",
         "rft_diffuse", "film_centre", "rr_start_weight"]


def set_slip(name):
    lib().ho_set_slip(SLIPS.index(name))


def film_splat(filter_type, radius, stddev, width, height, pos, val, b=1.0 / 3, c=1.0 / 3):
    """ImageBlock::Put of n samples (pos (n,2), val (n,3)) -> (H, W, 4) float64 sums."""
    pos = np.ascontiguousarray(pos, np.float32)
    val = np.ascontiguousarray(val, np.float32)
    out = np.zeros((height, width, 4), np.float64)
    _check(lib().ho_film_splat(filter_type, radius, stddev, b, c, width, height, pos.shape[0], _p(pos), _p(val),
                               _p(out)))
    return out


def fresnel_dielectric(cos_i, eta):
    ct = C.c_double()
    f = lib().ho_fresnel_dielectric(cos_i, eta, C.byref(ct))
    return f, ct.value


def fresnel_conductor(cos_i, eta, k):
    e = np.asarray(eta, np.float64)
    kk = np.asarray(k, np.float64)
    out = np.zeros(3)
    lib().ho_fresnel_conductor(cos_i, _p(e), _p(kk), _p(out))
    return out


def microfacet(dist, au, av, m, v, u):
    m = np.ascontiguousarray(m, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    n = m.shape[0]
    d, g1, sp = np.zeros(n), np.zeros(n), np.zeros(n)
    s = np.zeros((n, 3))
    lib().ho_microfacet(dist, au, av, n, _p(m), _p(v), _p(d), _p(g1), _p(u), _p(s), _p(sp))
    return d, g1, s, sp


def render(scene, req_seed=0, spp=(0, 1), threads=None, fp64=False):
    """Returns (rgb (H,W,3) float32, acc (H,W,4) float64, stats [ext, shadow, paths, render_us])."""
    desc = scene.to_c()
    W, H = scene.width, scene.height
    rgb = np.zeros((H, W, 3), np.float32)
    acc = np.zeros((H, W, 4), np.float64)
    st = np.zeros(4, np.uint64)
    threads = threads or min(16, os.cpu_count() or 1)
    _check(lib().ho_render(int(fp64), C.byref(desc), req_seed, spp[0], spp[1], threads, _p(rgb), _p(acc), _p(st)))
    return rgb, acc, st


def pixel_samples(scene, x, y, s0, n, req_seed=0, fp64=False):
    desc = scene.to_c()
    out = np.zeros((n, 5), np.float32)
    _check(lib().ho_pixel_samples(int(fp64), C.byref(desc), req_seed, x, y, s0, n, _p(out)))
    return out
