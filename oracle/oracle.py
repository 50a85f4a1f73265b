"""ctypes front end of the CPU oracle (rx_oracle.cpp) plus scene packing.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()``
and bench.py's CPU-baseline leg, never by the product package.  The BVH
the oracle traverses is built by bvh_oracle (the numpy restatement of the
reference builder), not by the GPU library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from . import bvh_oracle

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int32)


class OrcScene(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("aabb_min", fp), ("aabb_max", fp), ("left", ip), ("right", ip),
                ("tri_start", ip), ("tri_count", ip), ("n_tris", C.c_int), ("v0", fp), ("v1", fp), ("v2", fp),
                ("n0", fp), ("n1", fp), ("n2", fp), ("mat_idx", ip), ("mat_type", ip), ("mat_albedo", fp),
                ("mat_roughness", fp), ("mat_ior", fp), ("mat_emission", fp), ("n_prims", C.c_int),
                ("obj_type", ip), ("obj_center", fp), ("obj_radius", fp), ("obj_normal", fp), ("obj_offset", fp),
                ("obj_extra", fp), ("obj_v0", fp), ("obj_v1", fp), ("obj_v2", fp), ("obj_mat_type", ip),
                ("obj_albedo", fp), ("obj_roughness", fp), ("obj_ior", fp), ("obj_emission", fp),
                ("n_lights", C.c_int), ("light_center", fp), ("light_radius", fp), ("light_albedo", fp),
                ("light_emission", fp), ("mat_texture", ip), ("has_uv", ip), ("uv0", fp), ("uv1", fp), ("uv2", fp),
                ("n_textures", C.c_int), ("tex_width", ip), ("tex_height", ip), ("tex_flip_v", ip),
                ("tex_offset", C.POINTER(C.c_longlong)), ("texels", C.POINTER(C.c_ubyte))]


class OrcCamera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_tan", C.c_float), ("aspect", C.c_float), ("width", C.c_int32), ("height", C.c_int32)]


class OrcParams(C.Structure):
    _fields_ = [("max_depth", C.c_int32), ("use_sky", C.c_int32), ("direct_light_mode", C.c_int32),
                ("direct_light_max_depth", C.c_int32), ("sample_clamp", C.c_float), ("bg", C.c_float * 3),
                ("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32)]


class OrcCounts(C.Structure):
    _fields_ = [(n, C.c_ulonglong) for n in ("rays", "shadow_rays", "node_pops", "tri_tests", "prim_tests",
                                              "shadow_node_pops", "shadow_tri_tests", "shadow_prim_tests")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        h.orc_render_f32.restype = None
        h.orc_render_f64.restype = None
        h.orc_render_f64_moments.restype = None
        _lib = h
    return _lib


def _f(a):
    return a.ctypes.data_as(fp)


def _i(a):
    return a.ctypes.data_as(ip)


def threads_default():
    return int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))


class OracleScene:
    """Scene in the reference field layout, BVH from bvh_oracle."""

    def __init__(self, v0, v1, v2, n0, n1, n2, tri_mat, palette, prims=None, builder="morton", max_leaf=8,
                 uvs=None, has_uv=None, textures=None):
        """uvs: (uv0, uv1, uv2) per triangle (n x 2 each); has_uv per triangle;
        textures: list of (u8 HxWx3 texels, flip_v) indexed by the palette's
        texture column."""
        self.keep = []
        tri_mat = np.asarray(tri_mat, np.int32)
        n = len(tri_mat)
        if n:
            build = {"morton": bvh_oracle.build_morton, "median": bvh_oracle.build_median,
                     "object_median": bvh_oracle.build_object_median}[builder]
            self.tree = build(v0, v1, v2, max_leaf)
        else:
            self.tree = bvh_oracle.Tree(0)
        o = self.tree.order

        def g(a, shape=(-1, 3)):
            return np.ascontiguousarray(np.asarray(a, np.float32).reshape(shape)[o]) if n else np.zeros(
                (1, 3), np.float32)

        self.v0, self.v1, self.v2 = g(v0), g(v1), g(v2)
        self.n0, self.n1, self.n2 = g(n0), g(n1), g(n2)
        self.mat_idx = np.ascontiguousarray(tri_mat[o]) if n else np.zeros(1, np.int32)
        pal = list(palette) or [(0, (0.0, 0.0, 0.0), 0.0, 1.0, 0.0, -1)]
        self.mat_type = np.asarray([p[0] for p in pal], np.int32)
        self.mat_albedo = np.asarray([p[1] for p in pal], np.float32).reshape(-1, 3)
        self.mat_roughness = np.asarray([p[2] for p in pal], np.float32)
        self.mat_ior = np.asarray([p[3] for p in pal], np.float32)
        self.mat_emission = np.asarray([p[4] for p in pal], np.float32)
        self.mat_texture = np.asarray([p[5] if len(p) > 5 else -1 for p in pal], np.int32)
        if uvs is not None and n:
            self.uv = [np.ascontiguousarray(np.asarray(u, np.float32).reshape(-1, 2)[o]) for u in uvs]
        else:
            self.uv = [np.zeros((max(n, 1), 2), np.float32) for _ in range(3)]
        self.has_uv = (np.ascontiguousarray(np.asarray(has_uv, np.int32)[o]) if has_uv is not None and n
                       else np.zeros(max(n, 1), np.int32))
        tex = list(textures or [])
        self.tex_width = np.asarray([t[0].shape[1] for t in tex] or [0], np.int32)
        self.tex_height = np.asarray([t[0].shape[0] for t in tex] or [0], np.int32)
        self.tex_flip = np.asarray([1 if t[1] else 0 for t in tex] or [0], np.int32)
        sizes = [t[0].size for t in tex]
        self.tex_offset = np.asarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]) if tex else [0], np.int64)
        self.texels = (np.ascontiguousarray(np.concatenate([np.asarray(t[0], np.uint8).reshape(-1) for t in tex]))
                       if tex else np.zeros(1, np.uint8))
        self.n_textures = len(tex)
        t = self.tree
        if t.n_nodes == 0:
            t = bvh_oracle.Tree(1)
        self.nodes = [np.ascontiguousarray(x) for x in (t.aabb_min, t.aabb_max, t.left, t.right, t.tri_start,
                                                          t.tri_count)]
        P = prims or {}
        k = len(P.get("obj_type", []))

        def pf(name, cols=3):
            a = np.asarray(P.get(name, np.zeros((k, cols) if cols > 1 else k)), np.float32)
            a = a.reshape((k, cols) if cols > 1 else (k,))
            return np.ascontiguousarray(a) if k else np.zeros((1, cols) if cols > 1 else 1, np.float32)

        self.obj = {
            "obj_type": np.ascontiguousarray(np.asarray(P.get("obj_type", []), np.int32)) if k else np.zeros(1, np.int32),
            "obj_center": pf("obj_center"), "obj_radius": pf("obj_radius", 1), "obj_normal": pf("obj_normal"),
            "obj_offset": pf("obj_offset", 1), "obj_extra": pf("obj_extra"), "obj_v0": pf("obj_v0"),
            "obj_v1": pf("obj_v1"), "obj_v2": pf("obj_v2"),
            "obj_mat_type": np.ascontiguousarray(np.asarray(P.get("obj_mat_type", []), np.int32)) if k else np.zeros(1, np.int32),
            "obj_albedo": pf("obj_albedo"), "obj_roughness": pf("obj_roughness", 1), "obj_ior": pf("obj_ior", 1),
            "obj_emission": pf("obj_emission", 1),
        }
        if k:
            sel = ((self.obj["obj_type"] == 0) & (self.obj["obj_mat_type"] == 3) & (self.obj["obj_emission"] > 0))
        else:
            sel = np.zeros(1, bool)
        nl = int(sel.sum())
        self.lights = {
            "light_center": np.ascontiguousarray(self.obj["obj_center"][sel]) if nl else np.zeros((1, 3), np.float32),
            "light_radius": np.ascontiguousarray(self.obj["obj_radius"][sel]) if nl else np.zeros(1, np.float32),
            "light_albedo": np.ascontiguousarray(self.obj["obj_albedo"][sel]) if nl else np.zeros((1, 3), np.float32),
            "light_emission": np.ascontiguousarray(self.obj["obj_emission"][sel]) if nl else np.zeros(1, np.float32),
        }
        self.s = OrcScene()
        s = self.s
        s.n_nodes = self.tree.n_nodes
        s.aabb_min, s.aabb_max = _f(self.nodes[0]), _f(self.nodes[1])
        s.left, s.right, s.tri_start, s.tri_count = (_i(x) for x in self.nodes[2:])
        s.n_tris = n
        s.v0, s.v1, s.v2, s.n0, s.n1, s.n2 = (_f(x) for x in (self.v0, self.v1, self.v2, self.n0, self.n1, self.n2))
        s.mat_idx = _i(self.mat_idx)
        s.mat_type = _i(self.mat_type)
        s.mat_albedo, s.mat_roughness, s.mat_ior, s.mat_emission = (
            _f(x) for x in (self.mat_albedo, self.mat_roughness, self.mat_ior, self.mat_emission))
        s.n_prims = k
        for name, arr in self.obj.items():
            setattr(s, name, _i(arr) if arr.dtype == np.int32 else _f(arr))
        s.n_lights = nl
        for name, arr in self.lights.items():
            setattr(s, name, _f(arr))
        s.mat_texture = _i(self.mat_texture)
        s.has_uv = _i(self.has_uv)
        s.uv0, s.uv1, s.uv2 = (_f(u) for u in self.uv)
        s.n_textures = self.n_textures
        s.tex_width, s.tex_height, s.tex_flip_v = _i(self.tex_width), _i(self.tex_height), _i(self.tex_flip)
        s.tex_offset = self.tex_offset.ctypes.data_as(C.POINTER(C.c_longlong))
        s.texels = self.texels.ctypes.data_as(C.POINTER(C.c_ubyte))


def closest_hit(sc: OracleScene, rays, threads=None):
    rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(6, -1))
    n = rays.shape[1]
    t = np.empty(n, np.float32)
    idx = np.empty(n, np.int32)
    kind = np.empty(n, np.int32)
    u = np.empty(n, np.float32)
    v = np.empty(n, np.float32)
    cnt = OrcCounts()
    lib().orc_closest_hit(C.byref(sc.s), _f(rays), n, _f(t), _i(idx), _i(kind), _f(u), _f(v), C.byref(cnt),
                          int(threads or threads_default()))
    return {"t": t, "id": idx, "kind": kind, "u": u, "v": v, "counts": cnt.as_dict()}


def any_hit(sc: OracleScene, rays, tmax, threads=None):
    rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(6, -1))
    tmax = np.ascontiguousarray(np.asarray(tmax, np.float32))
    n = rays.shape[1]
    occ = np.empty(n, np.int32)
    lib().orc_any_hit(C.byref(sc.s), _f(rays), _f(tmax), n, _i(occ), int(threads or threads_default()))
    return occ


def hit_normals(sc: OracleScene, rays, hit):
    rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(6, -1))
    n = rays.shape[1]
    out = np.empty((n, 3), np.float32)
    lib().orc_hit_normals(C.byref(sc.s), _f(rays), n, _f(np.ascontiguousarray(hit["t"])), _i(hit["id"]),
                          _i(hit["kind"]), _f(np.ascontiguousarray(hit["u"])), _f(np.ascontiguousarray(hit["v"])),
                          _f(out))
    return out


def camera(pos, fwd, right, up, fov_tan, aspect, width, height):
    return OrcCamera((C.c_float * 3)(*map(float, pos)), (C.c_float * 3)(*map(float, fwd)),
                     (C.c_float * 3)(*map(float, right)), (C.c_float * 3)(*map(float, up)), float(fov_tan),
                     float(aspect), int(width), int(height))


def params(max_depth=4, use_sky=0, direct_light_mode=0, direct_light_max_depth=1, sample_clamp=10.0,
           bg=(0.0, 0.0, 0.0), seed=0):
    seed = int(seed) & ((1 << 64) - 1)
    return OrcParams(int(max_depth), int(use_sky), int(direct_light_mode), int(direct_light_max_depth),
                     float(sample_clamp), (C.c_float * 3)(*map(float, bg)), seed & 0xFFFFFFFF, seed >> 32)


class Accum:
    """Running means over a pixel set (all pixels, or a given list)."""

    def __init__(self, n, precision="f32", moments=False):
        self.precision = precision
        self.color = np.zeros((n, 3), np.float64 if precision == "f64" else np.float32)
        # sum of squared (clamped) samples per pixel and channel, f64 only
        self.sq = np.zeros((n, 3), np.float64) if moments else None
        self.normal = np.zeros((n, 3), np.float32)
        self.albedo = np.zeros((n, 3), np.float32)
        self.depth = np.zeros(n, np.float32)
        self.counts = OrcCounts()
        self.samples = 0


def render(sc: OracleScene, cam: OrcCamera, prm: OrcParams, first, count, acc: Accum, threads=None, pixels=None):
    """Samples [first, first+count) folded into ``acc`` (trace.py:239-283)."""
    plist = None
    nl = 0
    if pixels is not None:
        plist = np.ascontiguousarray(np.asarray(pixels, np.int32))
        nl = len(plist)
    fn = lib().orc_render_f64 if acc.precision == "f64" else lib().orc_render_f32
    col = acc.color.ctypes.data_as(C.POINTER(C.c_double)) if acc.precision == "f64" else _f(acc.color)
    args = [C.byref(sc.s), C.byref(cam), C.byref(prm), int(first), int(count), col, _f(acc.normal), _f(acc.albedo),
            _f(acc.depth), C.byref(acc.counts), int(threads or threads_default()),
            _i(plist) if plist is not None else None, nl]
    if acc.sq is not None:
        if acc.precision != "f64":
            raise ValueError("second moments are kept by the fp64 instance only")
        fn = lib().orc_render_f64_moments
        args.append(acc.sq.ctypes.data_as(C.POINTER(C.c_double)))
    fn(*args)
    acc.samples = first + count
    return acc


def adaptive_errors(sc: OracleScene, cam: OrcCamera, prm: OrcParams, samples, acc: Accum, threads=None):
    """rendering/ray_tracer.py:497-541 without the stop: renders one sample
    at a time and returns e[n] (n = samples rendered, from 2) = mean over
    pixels and channels of |accum_n - accum_(n-1)|, the error the adaptive
    rule compares with its threshold (:534), in float64 of the accumulator."""
    errs = {}
    prev = None
    for frame in range(samples):
        render(sc, cam, prm, frame, 1, acc, threads)
        cur = np.asarray(acc.color, dtype=np.float64).copy()
        if prev is not None:
            errs[frame + 1] = float(np.mean(np.abs(cur - prev)))
        prev = cur
    return errs


def adaptive_stop(errs, samples, min_samples, threshold, interval):
    """The sample count at which rendering/ray_tracer.py:531-541 stops (or
    `samples`), given the errors e[n] of adaptive_errors."""
    for n in range(2, samples + 1):
        if n >= samples or n < min_samples or (n - min_samples) % interval:
            continue
        if errs[n] <= threshold:
            return n
    return samples


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    out = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, out)
    return [int(x) for x in out]


def sincos_turns(u):
    u = np.ascontiguousarray(np.asarray(u, np.float32))
    s = np.empty_like(u)
    c = np.empty_like(u)
    lib().orc_sincos_turns(_f(u), len(u), _f(s), _f(c))
    return s, c


def hash_sine(x):
    x = np.ascontiguousarray(np.asarray(x, np.float64))
    out = np.empty_like(x)
    lib().orc_hash_sine(x.ctypes.data_as(C.POINTER(C.c_double)), len(x), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def camera_rays(cam: OrcCamera, frame):
    out = np.empty((cam.height * cam.width, 3), np.float32)
    lib().orc_camera_rays(C.byref(cam), int(frame), _f(out))
    return out


SCATTER_KINDS = {"cosine": 0, "unit": 1, "metal": 2, "dielectric": 3}


def scatter_dirs(kind, u1, u2=None, d=(0.0, 0.0, -1.0), n=(0.0, 0.0, 1.0), param=0.0):
    """The oracle's scatter.py samplers on given uniforms (orc_scatter_dirs):
    kind "cosine" (normal n), "unit", "metal" (incoming d, normal n, fuzz
    param), "dielectric" (d, n, eta param, xi = u1).  Returns (count, 3)."""
    u1 = np.ascontiguousarray(np.asarray(u1, np.float32))
    u2 = np.ascontiguousarray(np.asarray(u1 if u2 is None else u2, np.float32))
    out = np.empty((len(u1), 3), np.float32)
    dv = (C.c_float * 3)(*map(float, d))
    nv = (C.c_float * 3)(*map(float, n))
    lib().orc_scatter_dirs(SCATTER_KINDS[kind], dv, nv, C.c_float(param), _f(u1), _f(u2), len(u1), _f(out))
    return out
