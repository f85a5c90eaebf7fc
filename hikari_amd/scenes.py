"""Seeded synthetic scenes of the benchmark configurations.

* cbox(): closed box of 12 inward-facing triangles (albedo 0.725, red and green
  side walls), a 2-triangle ceiling light of radiance 15 and two diffuse
  icospheres (subdivision 3, 1280 triangles each, radius 0.3); pinhole 45 deg.
* displaced_sphere(): icosphere of subdivision 8 (1,310,720 triangles) with
  vertex radii 1 + 0.08 * fBm (5 octaves, seeded), GGX rough conductor with
  RGB eta / k, diffuse ground and a 512 x 256 synthetic environment map
  (seeded sky gradient plus an 8-texel sun patch).
* heightfield(): seeded random 1024 x 1024 heightfield (2,093,058 triangles)
  split into 8 patches of 5 material kinds, two quad lights plus the same map.

All geometry winds so that front faces look at the camera / into the box:
the reference's BSDFs and area lights are one-sided.
"""
from __future__ import annotations

import math

import numpy as np

from . import desc as D


# ----------------------------------------------------------------- geometry
def icosphere(subdiv, radius=1.0, center=(0, 0, 0)):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // (len(v) + 1), uk % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1).T + len(v)
        v = np.concatenate([v, mid])
        f = np.concatenate([np.stack([f[:, 0], m[:, 0], m[:, 2]], 1), np.stack([f[:, 1], m[:, 1], m[:, 0]], 1),
                            np.stack([f[:, 2], m[:, 2], m[:, 1]], 1), m])
    return v * radius + np.asarray(center, np.float64), f


def _ensure_outward(v, f, center):
    p0, p1, p2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    c = (p0 + p1 + p2) / 3 - np.asarray(center)
    flip = (n * c).sum(1) < 0
    f = f.copy()
    f[flip] = f[flip][:, [0, 2, 1]]
    return f


def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2),(p0,p2,p3); front side follows (p1-p0)x(p2-p0)."""
    return np.array([p0, p1, p2, p3], np.float64), np.array([[0, 1, 2], [0, 2, 3]], np.int64)


def fbm3(p, seed, octaves=5, lacunarity=2.0, gain=0.5, freq=2.0):
    """Seeded fractal value noise on R^3 (trilinear value noise per octave)."""
    rng = np.random.default_rng(seed)
    table = rng.uniform(-1.0, 1.0, 4096)
    perm = rng.permutation(4096)
    out = np.zeros(len(p))
    amp, total = 1.0, 0.0
    for o in range(octaves):
        q = p * freq + o * 17.31
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3 - 2 * fr)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    h = perm[(perm[(perm[(i0[:, 0] + dx) & 4095] + i0[:, 1] + dy) & 4095] + i0[:, 2] + dz) & 4095]
                    wt = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * (
                        w[:, 2] if dz else 1 - w[:, 2])
                    acc += wt * table[h]
        out += amp * acc
        total += amp
        amp *= gain
        freq *= lacunarity
    return out / total


def env_map(seed=7, w=512, h=256):
    """Seeded sky gradient (zenith blue -> horizon white, dim ground) + an 8-texel sun patch."""
    rng = np.random.default_rng(seed)
    theta = (np.arange(h) + 0.5) / h * math.pi
    up = np.cos(theta)[:, None]
    sky = np.where(up > 0, (1 - up) * np.array([0.9, 0.9, 1.0]) + up * np.array([0.25, 0.45, 0.95]),
                   np.array([0.12, 0.10, 0.08]) * (1 + up))
    img = np.repeat(sky[:, None, :], w, axis=1) * 0.8
    img *= 1 + 0.05 * rng.standard_normal((h, w, 1))
    sx, sy = int(rng.integers(0, w - 4)), int(rng.integers(h // 8, h // 3))
    img[sy:sy + 2, sx:sx + 4] = np.array([800.0, 760.0, 700.0])
    return np.ascontiguousarray(np.clip(img, 0, None), np.float32)


# Env maps here are z-up in the map's frame (EnvironmentLight.cpp:66 spherical
# direction with theta from +z); this frame puts the map's zenith on world +y.
ENV_Y_UP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, -1, 0, 0], [0, 0, 0, 1]], np.float64)


# ----------------------------------------------------------------- scenes
def cbox(width=256, height=256, spp_hint=64, sphere_subdiv=3, filter=D.FILTER_GAUSSIAN):
    """Config-0 closed box (see module docstring)."""
    s = D.Scene(width=width, height=height, fov=45.0, max_depth=1000, filter=filter)
    white = s.add_material(D.diffuse(0.725))
    red = s.add_material(D.diffuse((0.63, 0.065, 0.05)))
    green = s.add_material(D.diffuse((0.14, 0.45, 0.09)))
    black = s.add_material(D.diffuse(0.0))
    L = 1.0
    c = np.array([[-L, -L, -L], [L, -L, -L], [L, L, -L], [-L, L, -L], [-L, -L, L], [L, -L, L], [L, L, L], [-L, L, L]],
                 np.float64)

    def face(a, b, cc, d):  # quad wound so that its normal points into the box
        return np.array([c[a], c[b], c[cc], c[d]])

    walls = [  # floor (y=-1), ceiling (y=+1), back (z=-1), left (x=-1), right (x=+1), front (z=+1)
        (face(0, 4, 5, 1), white), (face(3, 2, 6, 7), white), (face(0, 1, 2, 3), white),
        (face(0, 3, 7, 4), red), (face(1, 5, 6, 2), green), (face(4, 7, 6, 5), white)]
    for q, m in walls:
        v, f = q, np.array([[0, 1, 2], [0, 2, 3]])
        cen = np.zeros(3)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        if np.dot(n, cen - v[0]) < 0:
            f = f[:, [0, 2, 1]]
        s.add_mesh(v, f, m)
    # ceiling light facing down, slightly below the ceiling
    y = L - 1e-3
    lv = np.array([[-0.25, y, -0.25], [0.25, y, -0.25], [0.25, y, 0.25], [-0.25, y, 0.25]])
    lf = np.array([[0, 1, 2], [0, 2, 3]])
    if np.cross(lv[1] - lv[0], lv[2] - lv[0])[1] > 0:
        lf = lf[:, [0, 2, 1]]
    s.add_mesh(lv, lf, black, radiance=(15.0, 15.0, 15.0))
    for cen in ((-0.45, -0.7, -0.2), (0.4, -0.7, 0.3)):
        v, f = icosphere(sphere_subdiv, 0.3, cen)
        s.add_mesh(v, _ensure_outward(v, f, cen), white)
    # camera inside the front wall, looking at -z
    s.cam_to_world = D.lookat((0, 0, 0.99), (0, 0, -1), (0, 1, 0))
    return s


CONDUCTORS = {  # RGB eta, k of common metals, relative to air
    "Au": ((0.143, 0.374, 1.442), (3.983, 2.386, 1.603)),
    "Cu": ((0.200, 0.924, 1.102), (3.912, 2.452, 2.142)),
    "Ag": ((0.155, 0.117, 0.138), (4.828, 3.122, 2.147)),
}


def displaced_sphere(alpha=0.3, metal="Au", width=1024, height=768, subdiv=8, seed=1, env_seed=7):
    """Config-1 scene (see module docstring)."""
    s = D.Scene(width=width, height=height, fov=40.0, max_depth=1000)
    eta, k = CONDUCTORS[metal]
    # the reference's conductors take one relative eta (intIOR / extIOR) and RGB k
    metal_m = s.add_material(D.rough_conductor(alpha=alpha, dist=D.GGX, int_ior=float(np.mean(eta)), ext_ior=1.0,
                                               k=k, ks=1.0))
    ground = s.add_material(D.diffuse(0.5))
    v, f = icosphere(subdiv, 1.0)
    r = 1.0 + 0.08 * fbm3(v, seed)
    v = v * r[:, None]
    f = _ensure_outward(v, f, (0, 0, 0))
    s.add_mesh(v, f, metal_m)
    gv, gf = quad((-20, -1.1, -20), (-20, -1.1, 20), (20, -1.1, 20), (20, -1.1, -20))
    if np.cross(gv[1] - gv[0], gv[2] - gv[0])[1] < 0:
        gf = gf[:, [0, 2, 1]]
    s.add_mesh(gv, gf, ground)
    s.env = env_map(env_seed)
    s.env_to_world = ENV_Y_UP
    s.cam_to_world = D.lookat((0.0, 0.6, 3.6), (0, -0.1, 0), (0, 1, 0))
    return s


def heightfield(width=1280, height=720, n=1024, seed=3, env_seed=7):
    """Config-2 scene (see module docstring)."""
    rng = np.random.default_rng(seed)
    s = D.Scene(width=width, height=height, fov=50.0, max_depth=1000)
    kinds = []
    for _ in range(8):
        kind = int(rng.integers(0, 5))
        if kind == 0:
            m = D.rough_dielectric(alpha=float(rng.uniform(0.1, 0.5)), dist=D.BECKMANN, int_ior=1.5, ext_ior=1.0)
        elif kind == 1:
            m = D.rough_plastic(alpha=float(rng.choice([0.15, 0.3])), dist=D.GGX, kd=float(rng.uniform(0.2, 0.8)))
        elif kind == 2:
            base = s.add_material(D.rough_conductor(alpha=0.2, dist=D.GGX, k=CONDUCTORS["Cu"][1]))
            m = D.coating(base, int_ior=1.5)
        elif kind == 3:
            m = D.dielectric(int_ior=1.5, ext_ior=1.0)
        else:
            m = D.diffuse(float(rng.uniform(0.3, 0.8)))
        kinds.append(s.add_material(m))
    xs = np.linspace(-6, 6, n)
    zs = np.linspace(-6, 6, n)
    X, Z = np.meshgrid(xs, zs)
    Y = 0.6 * fbm3(np.stack([X.ravel() * 0.2, Z.ravel() * 0.2, np.zeros(X.size)], 1), seed, octaves=5).reshape(n, n)
    V = np.stack([X, Y, Z], -1).reshape(-1, 3)
    # 8 patches: 4 x 2 blocks of the grid
    for p in range(8):
        i0, i1 = (p % 4) * (n - 1) // 4, ((p % 4) + 1) * (n - 1) // 4
        j0, j1 = (p // 4) * (n - 1) // 2, ((p // 4) + 1) * (n - 1) // 2
        ii, jj = np.meshgrid(np.arange(i0, i1), np.arange(j0, j1), indexing="ij")
        a = (jj * n + ii).ravel()
        b, c, d = a + 1, a + n + 1, a + n
        f = np.concatenate([np.stack([a, d, c], 1), np.stack([a, c, b], 1)])
        used = np.unique(f)
        remap = np.full(n * n, -1, np.int64)
        remap[used] = np.arange(len(used))
        s.add_mesh(V[used], remap[f], kinds[p])
    black = s.add_material(D.diffuse(0.0))
    for (x0, z0) in ((-3.0, -2.0), (2.5, 1.5)):
        lv, lf = quad((x0, 3.0, z0), (x0 + 1.2, 3.0, z0), (x0 + 1.2, 3.0, z0 + 1.2), (x0, 3.0, z0 + 1.2))
        if np.cross(lv[1] - lv[0], lv[2] - lv[0])[1] > 0:
            lf = lf[:, [0, 2, 1]]
        s.add_mesh(lv, lf, black, radiance=(12.0, 12.0, 12.0))
    s.env = env_map(env_seed)
    s.env_to_world = ENV_Y_UP
    s.cam_to_world = D.lookat((0.0, 4.0, 8.0), (0, 0, 0), (0, 1, 0))
    return s
