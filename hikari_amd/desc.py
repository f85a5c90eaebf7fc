"""Flat scene description shared by the C-ABI (include/hikari_amd.h) and tests.

A scene is a list of triangle meshes (float32 vertices, optional normals and
uv, uint32 faces), a table of material records (32 float32 each), an optional
lat-long environment map, a perspective camera, the reconstruction filter and
the pathMIS integrator settings.  The fields mirror the XML properties of the
reference plugins (include/core/Common.hpp:66-312 names, :314-465 defaults).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# Material type ids (record slot 0).  Parameter slots per type:
#   [1] distribution (0 beckmann, 1 ggx)
#   [2] nested material index (coatings; the upper side of a twoSided)
#   [3] flags (1 nonlinear, 2 anisotropic, 4 fastApprox) [4:7] colour A  [8:11] colour B
#   [12] lower-side material of a twoSided (-1: the upper one)
#   [20] intIOR [21] extIOR [22] alpha / alphaU [23] alphaV [24] thickness
# Colour A / B per type:
#   diffuse: A albedo | conductor, roughConductor: A ks, B k
#   dielectric, roughDielectric: A ksReflect, B ksRefract | plastic, roughPlastic: A ks, B kd
#   coating, roughCoating: A sigmaA, B ks | microfacet: A kd | roughDiffuse: A albedo
(DIFFUSE, CONDUCTOR, DIELECTRIC, PLASTIC, ROUGH_CONDUCTOR, ROUGH_DIELECTRIC, ROUGH_PLASTIC, COATING, ROUGH_COATING,
 MIRROR, MICROFACET, ROUGH_DIFFUSE, TWO_SIDED) = range(13)
BECKMANN, GGX = 0, 1
MAT_STRIDE = 32
AIR = 1.000277


def _rec(t, a=(0, 0, 0), b=(0, 0, 0), int_ior=0.0, ext_ior=0.0, alpha_u=0.0, alpha_v=0.0, thickness=0.0,
         dist=BECKMANN, nested=-1, flags=0, nested_back=-1):
    r = np.zeros(MAT_STRIDE, np.float32)
    r[0], r[1], r[2], r[3] = t, dist, nested, flags
    r[12] = nested_back
    r[4:7] = np.broadcast_to(np.asarray(a, np.float32), 3)
    r[8:11] = np.broadcast_to(np.asarray(b, np.float32), 3)
    r[20], r[21], r[22], r[23], r[24] = int_ior, ext_ior, alpha_u, alpha_v, thickness
    return r


def diffuse(albedo=0.5):
    return _rec(DIFFUSE, a=albedo)


def conductor(int_ior=1.5046, ext_ior=AIR, k=1.0, ks=1.0):
    return _rec(CONDUCTOR, a=ks, b=k, int_ior=int_ior, ext_ior=ext_ior)


def dielectric(int_ior=1.5046, ext_ior=AIR, ks_reflect=1.0, ks_refract=1.0):
    return _rec(DIELECTRIC, a=ks_reflect, b=ks_refract, int_ior=int_ior, ext_ior=ext_ior)


def plastic(int_ior=1.5046, ext_ior=AIR, ks=1.0, kd=0.5, nonlinear=False):
    return _rec(PLASTIC, a=ks, b=kd, int_ior=int_ior, ext_ior=ext_ior, flags=int(nonlinear))


def rough_conductor(alpha=0.1, dist=BECKMANN, int_ior=1.5046, ext_ior=AIR, k=1.0, ks=1.0, alpha_v=None):
    aniso = alpha_v is not None
    return _rec(ROUGH_CONDUCTOR, a=ks, b=k, int_ior=int_ior, ext_ior=ext_ior, alpha_u=alpha,
                alpha_v=alpha if alpha_v is None else alpha_v, dist=dist, flags=2 if aniso else 0)


def rough_dielectric(alpha=0.1, dist=BECKMANN, int_ior=1.5046, ext_ior=AIR, ks_reflect=1.0, ks_refract=1.0,
                     alpha_v=None):
    aniso = alpha_v is not None
    return _rec(ROUGH_DIELECTRIC, a=ks_reflect, b=ks_refract, int_ior=int_ior, ext_ior=ext_ior, alpha_u=alpha,
                alpha_v=alpha if alpha_v is None else alpha_v, dist=dist, flags=2 if aniso else 0)


def rough_plastic(alpha=0.1, dist=BECKMANN, int_ior=1.49, ext_ior=AIR, ks=1.0, kd=0.5, nonlinear=False):
    return _rec(ROUGH_PLASTIC, a=ks, b=kd, int_ior=int_ior, ext_ior=ext_ior, alpha_u=alpha, alpha_v=alpha,
                dist=dist, flags=int(nonlinear))


def coating(nested, int_ior=1.49, ext_ior=AIR, thickness=1.0, sigma_a=0.0, ks=1.0):
    return _rec(COATING, a=sigma_a, b=ks, int_ior=int_ior, ext_ior=ext_ior, thickness=thickness, nested=nested)


def rough_coating(nested, alpha=0.1, dist=BECKMANN, int_ior=1.49, ext_ior=AIR, thickness=1.0, sigma_a=0.0, ks=1.0):
    return _rec(ROUGH_COATING, a=sigma_a, b=ks, int_ior=int_ior, ext_ior=ext_ior, alpha_u=alpha, alpha_v=alpha,
                thickness=thickness, dist=dist, nested=nested)


def mirror():
    return _rec(MIRROR)


def microfacet(alpha=0.1, int_ior=1.5046, ext_ior=AIR, kd=0.5):
    return _rec(MICROFACET, a=kd, int_ior=int_ior, ext_ior=ext_ior, alpha_u=alpha, alpha_v=alpha)


def rough_diffuse(albedo=0.5, alpha=0.2, fast_approx=False):
    return _rec(ROUGH_DIFFUSE, a=albedo, alpha_u=alpha, alpha_v=alpha, flags=4 if fast_approx else 0)


def two_sided(front, back=-1):
    return _rec(TWO_SIDED, nested=front, nested_back=back)


FILTER_BOX, FILTER_GAUSSIAN, FILTER_TENT, FILTER_MITCHELL = 0, 1, 2, 3


class MeshDesc(C.Structure):
    _fields_ = [("v", C.POINTER(C.c_float)), ("nv", C.c_uint32),
                ("n", C.POINTER(C.c_float)), ("uv", C.POINTER(C.c_float)),
                ("f", C.POINTER(C.c_uint32)), ("nt", C.c_uint32),
                ("material", C.c_int32), ("is_emitter", C.c_int32), ("radiance", C.c_float * 3)]


class SceneDesc(C.Structure):
    _fields_ = [("meshes", C.POINTER(MeshDesc)), ("n_meshes", C.c_int32),
                ("materials", C.POINTER(C.c_float)), ("n_materials", C.c_int32),
                ("env", C.POINTER(C.c_float)), ("env_w", C.c_int32), ("env_h", C.c_int32),
                ("env_scale", C.c_float), ("env_to_world", C.c_float * 16),
                ("cam_to_world", C.c_float * 16), ("fov", C.c_float), ("near_clip", C.c_float),
                ("far_clip", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("filter", C.c_int32), ("filter_radius", C.c_float), ("filter_stddev", C.c_float),
                ("filter_b", C.c_float), ("filter_c", C.c_float),
                ("max_depth", C.c_int32), ("background", C.c_float * 3), ("force_background", C.c_int32),
                ("rft_dir", C.c_char_p), ("integrator", C.c_int32), ("rr_start_depth", C.c_int32)]

INTEGRATORS = {"pathMIS": 0, "pathEMS": 1, "pathMATS": 2, "whitted": 3}


def lookat(origin, target, up):
    """Camera-to-world matrix of the XML <lookat> (src/core/Parser.cpp:349-361)."""
    o, t, u = (np.asarray(x, np.float64) for x in (origin, target, up))
    d = t - o
    d /= np.linalg.norm(d)
    left = np.cross(u / np.linalg.norm(u), d)
    left /= np.linalg.norm(left)
    nup = np.cross(d, left)
    nup /= np.linalg.norm(nup)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = left, nup, d, o
    return m


DATA_DIR = __import__("os").path.join(__import__("os").path.dirname(__file__), "data")


@dataclass
class Mesh:
    v: np.ndarray
    f: np.ndarray
    material: int = 0
    n: Optional[np.ndarray] = None
    uv: Optional[np.ndarray] = None
    radiance: Optional[tuple] = None


@dataclass
class Scene:
    meshes: List[Mesh] = field(default_factory=list)
    materials: List[np.ndarray] = field(default_factory=list)
    env: Optional[np.ndarray] = None  # (h, w, 3) float32
    env_scale: float = 1.0
    env_to_world: np.ndarray = field(default_factory=lambda: np.eye(4))
    cam_to_world: np.ndarray = field(default_factory=lambda: np.eye(4))
    fov: float = 30.0
    near_clip: float = 1e-4
    far_clip: float = 1e4
    width: int = 1280
    height: int = 720
    filter: int = FILTER_GAUSSIAN
    filter_radius: float = 2.0
    filter_stddev: float = 0.5
    filter_b: float = 1.0 / 3.0
    filter_c: float = 1.0 / 3.0
    max_depth: int = 1000
    background: tuple = (0.0, 0.0, 0.0)
    force_background: bool = False
    integrator: str = "pathMIS"
    rr_start_depth: int = 0

    def add_material(self, rec) -> int:
        self.materials.append(np.asarray(rec, np.float32))
        return len(self.materials) - 1

    def add_mesh(self, v, f, material, n=None, uv=None, radiance=None):
        self.meshes.append(Mesh(np.ascontiguousarray(v, np.float32), np.ascontiguousarray(f, np.uint32), material,
                                None if n is None else np.ascontiguousarray(n, np.float32),
                                None if uv is None else np.ascontiguousarray(uv, np.float32), radiance))

    @property
    def n_triangles(self):
        return sum(int(m.f.shape[0]) for m in self.meshes)

    def to_c(self):
        """Build the ctypes SceneDesc; the returned object keeps the arrays alive."""
        keep = []
        fp = C.POINTER(C.c_float)
        arr = (MeshDesc * max(1, len(self.meshes)))()
        for i, m in enumerate(self.meshes):
            keep += [m.v, m.f]
            d = arr[i]
            d.v = m.v.ctypes.data_as(fp)
            d.nv = m.v.shape[0]
            d.f = m.f.ctypes.data_as(C.POINTER(C.c_uint32))
            d.nt = m.f.shape[0]
            if m.n is not None:
                keep.append(m.n)
                d.n = m.n.ctypes.data_as(fp)
            if m.uv is not None:
                keep.append(m.uv)
                d.uv = m.uv.ctypes.data_as(fp)
            d.material = m.material
            d.is_emitter = int(m.radiance is not None)
            if m.radiance is not None:
                d.radiance[:] = [float(x) for x in m.radiance]
        mats = np.ascontiguousarray(np.stack(self.materials) if self.materials else np.zeros((1, MAT_STRIDE)),
                                    np.float32)
        keep += [arr, mats]
        s = SceneDesc()
        s.meshes = C.cast(arr, C.POINTER(MeshDesc))
        s.n_meshes = len(self.meshes)
        s.materials = mats.ctypes.data_as(fp)
        s.n_materials = len(self.materials)
        if self.env is not None:
            env = np.ascontiguousarray(self.env, np.float32)
            keep.append(env)
            s.env = env.ctypes.data_as(fp)
            s.env_h, s.env_w = env.shape[0], env.shape[1]
        s.env_scale = self.env_scale
        s.env_to_world[:] = [float(x) for x in np.asarray(self.env_to_world, np.float64).ravel()]
        s.cam_to_world[:] = [float(x) for x in np.asarray(self.cam_to_world, np.float64).ravel()]
        s.fov, s.near_clip, s.far_clip = self.fov, self.near_clip, self.far_clip
        s.width, s.height = self.width, self.height
        s.filter, s.filter_radius, s.filter_stddev = self.filter, self.filter_radius, self.filter_stddev
        s.filter_b, s.filter_c = self.filter_b, self.filter_c
        s.max_depth = self.max_depth
        s.background[:] = [float(x) for x in self.background]
        s.force_background = int(self.force_background)
        s.integrator = INTEGRATORS[self.integrator or "pathMIS"]
        s.rr_start_depth = int(self.rr_start_depth)
        rft = DATA_DIR.encode()
        keep.append(rft)
        s.rft_dir = rft
        s._keep = keep  # noqa: attribute on ctypes struct instance keeps buffers alive
        return s


def load_obj(path, to_world=None):
    """Wavefront OBJ reader following src/mesh/WavefrontObjMesh.cpp:10-152.

    Quads are split (0,1,2),(3,0,2); vertices are unique per (p, uv, n) index
    triple in first-use order; positions go through to_world with the
    homogeneous divide, normals through its inverse transpose and are
    re-normalised.
    """
    P, T, N, idx, verts, vmap = [], [], [], [], [], {}
    M = np.eye(4) if to_world is None else np.asarray(to_world, np.float64)
    Minv_t = np.linalg.inv(M)[:3, :3].T
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                p = np.array([float(x) for x in tok[1:4]] + [1.0], np.float32).astype(np.float64)
                q = M @ p
                P.append((q[:3] / q[3]).astype(np.float32))
            elif tok[0] == "vt":
                T.append(np.array([float(x) for x in tok[1:3]], np.float32))
            elif tok[0] == "vn":
                n = Minv_t @ np.array([float(x) for x in tok[1:4]], np.float32).astype(np.float64)
                n = n.astype(np.float32)
                N.append(n / np.float32(math.sqrt(float(n[0] * n[0] + (n[1] * n[1] + n[2] * n[2])))))
            elif tok[0] == "f":
                vs = tok[1:5]
                order = [0, 1, 2] if len(vs) == 3 else [0, 1, 2, 3, 0, 2]
                for k in order:
                    parts = vs[k].split("/")
                    key = (int(parts[0]), int(parts[1]) if len(parts) > 1 and parts[1] else 0,
                           int(parts[2]) if len(parts) > 2 and parts[2] else 0)
                    if key not in vmap:
                        vmap[key] = len(verts)
                        verts.append(key)
                    idx.append(vmap[key])
    v = np.array([P[k[0] - 1] for k in verts], np.float32)
    uv = np.array([T[k[1] - 1] for k in verts], np.float32) if T else None
    n = np.array([N[k[2] - 1] for k in verts], np.float32) if N else None
    f = np.array(idx, np.uint32).reshape(-1, 3)
    return v, f, n, uv
