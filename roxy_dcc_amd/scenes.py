"""Seeded synthetic scenes for the benchmark configurations.

cornell(seed)          256^2 Cornell box: 5 walls (10 triangles), a 0.25^2
                       emissive ceiling quad at y = 0.999, two white boxes
                       (0.3^3 and 0.3x0.6x0.3) rotated -+18 deg about y.
mesh_scene(seed)       the Cornell enclosure around a displaced geodesic
                       icosphere (frequency 59: 69,620 triangles, 5% radial
                       noise) with a lambertian/metal/dielectric material,
                       and one emissive-sphere light; 512^2.
large_scene(seed)      500 displaced UV spheres (4,000 triangles each),
                       64 analytic spheres, three emissive-sphere lights in
                       a 100^3 box; 1920x1080.
batch_scene(seed)      one scene of the many-scene batch: a displaced UV
                       sphere of log-uniform [1k, 200k] triangles in the
                       Cornell enclosure.

Each returns (world, width, height, render settings).  Geometry goes in
as IndexedMesh objects, so flattening takes the vectorised path.
"""
from __future__ import annotations

import math

import numpy as np

from .world import (Camera, Cube, Dielectric, Diffuse, Emissive, IndexedMesh, Mat4, Metal, SceneObject, Shape,
                    Sphere, World)

RED = (0.65, 0.05, 0.05)
GREEN = (0.12, 0.45, 0.15)
WHITE = (0.73, 0.73, 0.73)


def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3)."""
    pos = np.array([p0, p1, p2, p3], dtype=np.float64)
    return pos, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)


def _mesh_object(name, pos, idx, material, transform=None):
    return SceneObject(name, [Shape(IndexedMesh(pos, idx), material)], transform)


def _box_mesh(sx, sy, sz):
    """Axis-aligned box [0,sx]x[0,sy]x[0,sz] as 12 outward-wound triangles."""
    p = np.array([[x, y, z] for x in (0, sx) for y in (0, sy) for z in (0, sz)], dtype=np.float64)
    # vertex index = 4*ix + 2*iy + iz
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, c, d in faces:
        tris += [[a, b, c], [a, c, d]]
    return p, np.array(tris, dtype=np.int32)


def cornell_objects():
    """Walls, ceiling light and the two boxes of the Cornell configuration."""
    objs = []
    walls = [
        ("left", ((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)), RED),
        ("right", ((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)), GREEN),
        ("floor", ((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)), WHITE),
        ("ceiling", ((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)), WHITE),
        ("back", ((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)), WHITE),
    ]
    for name, quad, alb in walls:
        pos, idx = _quad(*quad)
        objs.append(_mesh_object(name, pos, idx, Diffuse(alb)))
    lo, hi = 0.375, 0.625
    pos, idx = _quad((lo, 0.999, lo), (hi, 0.999, lo), (hi, 0.999, hi), (lo, 0.999, hi))
    objs.append(_mesh_object("light", pos, idx, Emissive((1.0, 1.0, 1.0), 15.0)))
    for name, size, at, angle in (("short", (0.3, 0.3, 0.3), (0.55, 0.0, 0.2), 18.0),
                                  ("tall", (0.3, 0.6, 0.3), (0.15, 0.0, 0.5), -18.0)):
        pos, idx = _box_mesh(*size)
        xf = Mat4.trs(translation=at, rotation_deg=(0.0, angle, 0.0))
        objs.append(_mesh_object(name, pos, idx, Diffuse(WHITE), xf))
    return objs


def _camera(world, pos, fwd, fov, w, h):
    world.add_camera(Camera(position=pos, forward=fwd, up=(0, 1, 0), fov=fov, width=w, height=h))


def cornell(seed=1):
    w = World(background_color=(0.0, 0.0, 0.0), use_sky=False)
    for o in cornell_objects():
        w.add_object(o)
    _camera(w, (0.5, 0.5, -1.4), (0.0, 0.0, 1.0), 40.0, 256, 256)
    return w, 256, 256, dict(samples=16, max_depth=5, direct_light_mode="one", direct_light_max_depth=1)


# ------------------------------------------------------------------ meshes
def _value_noise(rng, pts, cells=6):
    """Seeded trilinear value noise in [-1, 1] on the unit sphere's box."""
    lattice = rng.uniform(-1.0, 1.0, size=(cells + 1,) * 3)
    g = (np.clip(pts, -1.0, 1.0) + 1.0) * 0.5 * (cells - 1e-9)
    i0 = np.floor(g).astype(np.int64)
    f = g - i0
    f = f * f * (3 - 2 * f)
    out = np.zeros(len(pts))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wgt = ((f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1])
                       * (f[:, 2] if dz else 1 - f[:, 2]))
                out += wgt * lattice[i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz]
    return out


def geodesic_sphere(freq):
    """Icosahedron with each face split into freq^2 triangles, projected to
    the unit sphere (20 * freq^2 triangles; shared vertices per face)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    positions, tris = [], []
    for a, b, c in f:
        base = len(positions)
        ids = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                k = freq - i - j
                p = (v[a] * k + v[b] * i + v[c] * j) / freq
                ids[(i, j)] = len(positions)
                positions.append(p)
        for i in range(freq):
            for j in range(freq - i):
                tris.append([ids[(i, j)], ids[(i + 1, j)], ids[(i, j + 1)]])
                if j + i + 1 < freq:
                    tris.append([ids[(i + 1, j)], ids[(i + 1, j + 1)], ids[(i, j + 1)]])
        del base
    pos = np.asarray(positions)
    pos /= np.linalg.norm(pos, axis=1, keepdims=True)
    return pos, np.asarray(tris, dtype=np.int32)


def uv_sphere(seg_u, seg_v):
    """UV sphere with seg_u longitudes and seg_v latitude bands
    (2 * seg_u * (seg_v - 1) triangles; poles as triangle fans)."""
    th = np.pi * np.arange(1, seg_v) / seg_v
    ph = 2 * np.pi * np.arange(seg_u) / seg_u
    ring = np.stack([np.outer(np.sin(th), np.cos(ph)), np.repeat(np.cos(th)[:, None], seg_u, axis=1),
                     np.outer(np.sin(th), np.sin(ph))], axis=-1).reshape(-1, 3)
    pos = np.concatenate([[[0.0, 1.0, 0.0]], ring, [[0.0, -1.0, 0.0]]])
    south = len(pos) - 1
    j = np.arange(seg_u)
    jn = (j + 1) % seg_u

    def rid(i, jj):
        return 1 + (i - 1) * seg_u + jj

    north = np.stack([np.zeros(seg_u, np.int64), rid(1, jn), rid(1, j)], axis=1)
    i = np.arange(1, seg_v - 1)[:, None]
    a, b, c, d = rid(i, j), rid(i, jn), rid(i + 1, jn), rid(i + 1, j)
    band = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], axis=2).reshape(-1, 3)
    southf = np.stack([np.full(seg_u, south), rid(seg_v - 1, j), rid(seg_v - 1, jn)], axis=1)
    return pos, np.concatenate([north, band, southf]).astype(np.int32)


def displaced(pos, rng, amplitude=0.05):
    return pos * (1.0 + amplitude * _value_noise(rng, pos))[:, None]


def _random_material(rng):
    kind = rng.integers(0, 3)
    if kind == 0:
        return Diffuse(tuple(rng.uniform(0.2, 0.9, size=3)))
    if kind == 1:
        return Metal(tuple(rng.uniform(0.2, 0.9, size=3)), roughness=float(rng.uniform(0.0, 0.3)))
    return Dielectric(ior=1.5)


def mesh_scene(seed=2, freq=59):
    rng = np.random.default_rng(seed)
    w = World(background_color=(0.0, 0.0, 0.0), use_sky=False)
    for o in cornell_objects()[:5]:
        w.add_object(o)
    pos, idx = geodesic_sphere(freq)
    pos = displaced(pos, rng) * 0.25 + np.array([0.5, 0.3, 0.5])
    w.add_object(_mesh_object("blob", pos, idx, _random_material(rng)))
    w.add_object(SceneObject("light", [Shape(Sphere(0.08), Emissive((1.0, 1.0, 1.0), 15.0))],
                             Mat4.trs(translation=(0.5, 0.85, 0.5))))
    _camera(w, (0.5, 0.5, -1.4), (0.0, 0.0, 1.0), 40.0, 512, 512)
    return w, 512, 512, dict(samples=64, max_depth=8, direct_light_mode="one", direct_light_max_depth=1,
                             denoise_passes=5)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def large_scene(seed=3, n_meshes=500, seg_u=50, seg_v=41, n_spheres=64):
    rng = np.random.default_rng(seed)
    w = World(background_color=(0.0, 0.0, 0.0), use_sky=True)
    base_pos, base_idx = uv_sphere(seg_u, seg_v)
    for i in range(n_meshes):
        pos = displaced(base_pos, rng)
        m = np.eye(4)
        m[:3, :3] = _random_rotation(rng) * rng.uniform(0.5, 3.0)
        m[:3, 3] = rng.uniform(0.0, 100.0, size=3)
        w.add_object(_mesh_object(f"mesh{i}", pos, base_idx, _random_material(rng), Mat4(m)))
    for i in range(n_spheres):
        r = float(rng.uniform(0.5, 2.0))
        w.add_object(SceneObject(f"sphere{i}", [Shape(Sphere(r), _random_material(rng))],
                                 Mat4.trs(translation=tuple(rng.uniform(0.0, 100.0, size=3)))))
    for i in range(3):
        w.add_object(SceneObject(f"light{i}", [Shape(Sphere(4.0), Emissive((1.0, 1.0, 1.0), 20.0))],
                                 Mat4.trs(translation=tuple(rng.uniform(10.0, 90.0, size=3)))))
    _camera(w, (50.0, 50.0, -120.0), (0.0, 0.0, 1.0), 45.0, 1920, 1080)
    return w, 1920, 1080, dict(samples=16, max_depth=6, direct_light_mode="one", direct_light_max_depth=1)


def batch_scene(seed):
    """One job of the many-scene batch (seed 4000 + i)."""
    rng = np.random.default_rng(seed)
    n_tris = int(round(math.exp(rng.uniform(math.log(1000), math.log(200000)))))
    # UV sphere with 2*u*(v-1) ~ n_tris triangles and u ~ 2v
    v = max(3, int(round(math.sqrt(n_tris / 4.0))))
    u = max(3, int(round(n_tris / (2.0 * (v - 1)))))
    res = int(rng.choice([64, 128, 256, 512]))
    spp = int(rng.integers(4, 33))
    w = World(background_color=(0.0, 0.0, 0.0), use_sky=False)
    for o in cornell_objects()[:5]:
        w.add_object(o)
    pos, idx = uv_sphere(u, v)
    pos = displaced(pos, rng) * 0.25 + np.array([0.5, 0.3, 0.5])
    w.add_object(_mesh_object("blob", pos, idx, _random_material(rng)))
    w.add_object(SceneObject("light", [Shape(Sphere(0.08), Emissive((1.0, 1.0, 1.0), 15.0))],
                             Mat4.trs(translation=(0.5, 0.85, 0.5))))
    _camera(w, (0.5, 0.5, -1.4), (0.0, 0.0, 1.0), 40.0, res, res)
    return w, res, res, dict(samples=spp, max_depth=6, direct_light_mode="one", direct_light_max_depth=1)


CONFIGS = {"cornell": cornell, "mesh": mesh_scene, "large": large_scene}
