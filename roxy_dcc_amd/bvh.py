"""BVH construction behind the reference's builder interface.

Mirrors rendering/taichi/bvh_builder.py: ``TriangleBatch`` (:8-272) with
the same fields, coercions, validation errors and ``concat`` palette
offsetting, and ``GPUBVHBuilder`` (:275-655) with the same class knobs,
``build(tri_data, method=None)``, ``upload()`` and result attributes.

What changes is where the work runs.  ``build`` uploads the triangle
arrays once and runs the Morton build as gfx950 kernels
(``rx_bvh_build_morton``); the triangles are gathered into leaf order and
packed for traversal on the device (``rx_bvh_pack_triangles``).  The tree
stays resident in HBM: 64-byte two-child nodes, a 32-byte root header,
48-byte triangles (v0, e1, e2), 48-byte shading records and 32-byte
materials.  ``upload()`` therefore only checks the reference's capacity
limits.  The reference-layout node arrays (fields.py:50-56) are also kept
on the device and read back lazily for ``nodes``.

``method="median"`` (bvh_builder.py:488-525) keeps the reference's host
build, numpy ``argpartition`` included, because the order numpy gives the
elements inside each half is what defines that tree; its nodes are then
packed on the device with ``rx_bvh_pack``.

``method="object_median"`` (and ``build_object_median`` for fp64 vertex
arrays) is core/bvh.py's BVHNode rule — fp64 boxes and centroids, a stable
sort on the longest axis, a split at len // 2 — built on the GPU
(``rx_bvh_build_object_median``) and flattened in preorder into the same
layouts.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as nat

# capacity limits of the reference fields (rendering/taichi/fields.py:7-9)
MAX_BVH_NODES = 4_000_000
MAX_TRIANGLES = 2_000_000
MAX_MATERIALS = 4_096

MATERIAL_DTYPE = np.dtype([("type", "<i4"), ("albedo", "<f4", 3), ("roughness", "<f4"), ("ior", "<f4"),
                           ("emission", "<f4"), ("texture", "<i4")])
assert MATERIAL_DTYPE.itemsize == nat.MATERIAL_BYTES


def _vec3(value):
    return np.asarray(value, dtype=np.float32).reshape((-1, 3))


def _vec2(value, n):
    if value is None:
        return np.zeros((n, 2), dtype=np.float32)
    return np.asarray(value, dtype=np.float32).reshape((-1, 2))


@dataclass
class TriangleBatch:
    """Array-backed triangles (rendering/taichi/bvh_builder.py:8-272).

    Either per-triangle materials (``mat_idx`` None: mat_type/albedo/...
    have one row per triangle) or a compact palette (``mat_idx`` indexes
    rows of mat_type/albedo/..., ``mat_palette`` holds the same rows as
    dicts).
    """

    v0: np.ndarray
    v1: np.ndarray
    v2: np.ndarray
    n0: np.ndarray
    n1: np.ndarray
    n2: np.ndarray
    mat_type: np.ndarray
    albedo: np.ndarray
    roughness: np.ndarray
    ior: np.ndarray
    emission: np.ndarray
    uv0: np.ndarray = None
    uv1: np.ndarray = None
    uv2: np.ndarray = None
    has_uv: np.ndarray = None
    texture_idx: np.ndarray = None
    mat_idx: np.ndarray = None
    mat_palette: list = None

    def __post_init__(self):
        for name in ("v0", "v1", "v2", "n0", "n1", "n2", "albedo"):
            setattr(self, name, _vec3(getattr(self, name)))
        self.mat_type = np.asarray(self.mat_type, dtype=np.int32).reshape((-1,))
        for name in ("roughness", "ior", "emission"):
            setattr(self, name, np.asarray(getattr(self, name), dtype=np.float32).reshape((-1,)))
        n = self.triangle_count
        self.uv0, self.uv1, self.uv2 = _vec2(self.uv0, n), _vec2(self.uv1, n), _vec2(self.uv2, n)
        self.has_uv = (np.zeros(n, dtype=np.int32) if self.has_uv is None
                       else np.asarray(self.has_uv, dtype=np.int32).reshape((-1,)))
        n_tex = n if self.mat_idx is None else len(self.mat_type)
        self.texture_idx = (np.full(n_tex, -1, dtype=np.int32) if self.texture_idx is None
                            else np.asarray(self.texture_idx, dtype=np.int32).reshape((-1,)))

        geometry = (self.v0, self.v1, self.v2, self.n0, self.n1, self.n2, self.uv0, self.uv1, self.uv2, self.has_uv)
        if len({len(a) for a in geometry}) != 1:
            raise ValueError("TriangleBatch geometry arrays must all have the same length")
        material = (self.mat_type, self.albedo, self.roughness, self.ior, self.emission, self.texture_idx)
        lengths = {len(a) for a in material}
        if self.mat_idx is None:
            if len(lengths) != 1 or next(iter(lengths)) != n:
                raise ValueError("TriangleBatch per-triangle material arrays must match triangle count")
        else:
            self.mat_idx = np.asarray(self.mat_idx, dtype=np.int32).reshape((-1,))
            if len(self.mat_idx) != n:
                raise ValueError("TriangleBatch mat_idx must match triangle count")
            if len(lengths) != 1:
                raise ValueError("TriangleBatch material palette arrays must all have the same length")
            if self.mat_palette is None:
                self.mat_palette = [self._palette_row(i) for i in range(len(self.mat_type))]

    def _palette_row(self, i):
        return {
            "type": int(self.mat_type[i]),
            "albedo": self.albedo[i].tolist(),
            "roughness": float(self.roughness[i]),
            "ior": float(self.ior[i]),
            "emission": float(self.emission[i]),
            "texture_idx": int(self.texture_idx[i]),
        }

    @property
    def triangle_count(self):
        return len(self.v0)

    @classmethod
    def empty(cls):
        z3 = np.zeros((0, 3), dtype=np.float32)
        z = np.zeros(0, dtype=np.float32)
        zi = np.zeros(0, dtype=np.int32)
        z2 = np.zeros((0, 2), dtype=np.float32)
        return cls(z3, z3, z3, z3, z3, z3, zi, z3, z, z, z, uv0=z2, uv1=z2, uv2=z2, has_uv=zi, texture_idx=zi,
                   mat_idx=zi, mat_palette=[])

    @classmethod
    def from_dicts(cls, tris):
        """Legacy per-triangle dicts (bvh_builder.py:133-154)."""
        if not tris:
            return cls.empty()

        def col(key, default=None):
            return [t.get(key, default) if default is not None else t[key] for t in tris]

        return cls(col("v0"), col("v1"), col("v2"), col("n0"), col("n1"), col("n2"), col("mat_type"),
                   col("albedo"), col("roughness"), col("ior"), col("emission"),
                   uv0=col("uv0", [0.0, 0.0]), uv1=col("uv1", [0.0, 0.0]), uv2=col("uv2", [0.0, 0.0]),
                   has_uv=col("has_uv", 0), texture_idx=col("texture_idx", -1))

    @classmethod
    def from_indexed_arrays(cls, arrays, material_by_group_idx):
        """Indexed-mesh arrays with a compact palette of the groups actually
        used, in ascending group order (bvh_builder.py:156-212)."""
        group_idx = np.asarray(arrays["group_idx"], dtype=np.int32)
        used = np.unique(group_idx)
        if len(used):
            remap = np.zeros(int(used.max()) + 1, dtype=np.int32)
            remap[used] = np.arange(len(used), dtype=np.int32)
            local = remap[group_idx]
        else:
            local = group_idx
        k = len(used)
        mat_type = np.zeros(k, dtype=np.int32)
        albedo = np.zeros((k, 3), dtype=np.float32)
        rough = np.zeros(k, dtype=np.float32)
        ior = np.ones(k, dtype=np.float32)
        emis = np.zeros(k, dtype=np.float32)
        tex = np.full(k, -1, dtype=np.int32)
        for j, g in enumerate(used):
            m = material_by_group_idx[int(g)]
            mat_type[j] = m["mat_type"]
            albedo[j] = m["albedo"]
            rough[j] = m["roughness"]
            ior[j] = m["ior"]
            emis[j] = m["emission"]
            tex[j] = m.get("texture_idx", -1)
        batch = cls(arrays["v0"], arrays["v1"], arrays["v2"], arrays["n0"], arrays["n1"], arrays["n2"], mat_type,
                    albedo, rough, ior, emis, uv0=arrays.get("uv0"), uv1=arrays.get("uv1"), uv2=arrays.get("uv2"),
                    has_uv=arrays.get("has_uv"), texture_idx=tex, mat_idx=local)
        return batch

    @classmethod
    def concat(cls, batches):
        """Concatenate batches; palettes are appended with index offsets
        when every batch has one (bvh_builder.py:214-272)."""
        batches = [b for b in batches if b.triangle_count]
        if not batches:
            return cls.empty()
        if len(batches) == 1:
            return batches[0]

        def cat(name):
            return np.concatenate([getattr(b, name) for b in batches])

        geo = dict(uv0=cat("uv0"), uv1=cat("uv1"), uv2=cat("uv2"), has_uv=cat("has_uv"))
        if all(b.mat_idx is not None for b in batches):
            palette, idx = [], []
            for b in batches:
                idx.append(b.mat_idx + len(palette))
                palette.extend(b.mat_palette)
            return cls(cat("v0"), cat("v1"), cat("v2"), cat("n0"), cat("n1"), cat("n2"),
                       np.asarray([m["type"] for m in palette], dtype=np.int32),
                       np.asarray([m["albedo"] for m in palette], dtype=np.float32).reshape((-1, 3)),
                       np.asarray([m["roughness"] for m in palette], dtype=np.float32),
                       np.asarray([m["ior"] for m in palette], dtype=np.float32),
                       np.asarray([m["emission"] for m in palette], dtype=np.float32),
                       texture_idx=np.asarray([m.get("texture_idx", -1) for m in palette], dtype=np.int32),
                       mat_idx=np.concatenate(idx), mat_palette=palette, **geo)
        return cls(cat("v0"), cat("v1"), cat("v2"), cat("n0"), cat("n1"), cat("n2"), cat("mat_type"),
                   cat("albedo"), cat("roughness"), cat("ior"), cat("emission"), texture_idx=cat("texture_idx"),
                   **geo)


def material_key(mat):
    """Palette key of bvh_builder.py:389-399 (type, albedo, roughness, ior,
    emission, texture)."""
    mt = mat.get("mat_type", mat.get("type"))
    return (int(mt), tuple(float(v) for v in mat["albedo"]), float(mat["roughness"]), float(mat["ior"]),
            float(mat["emission"]), int(mat.get("texture_idx", -1)))


def morton_shape(n_tris: int, max_leaf: int) -> tuple[int, int]:
    """(node count, depth) of the Morton tree, from the library's closed form."""
    n, d = C.c_int(), C.c_int()
    nat.check(nat.lib().rx_bvh_morton_shape(int(n_tris), int(max_leaf), C.byref(n), C.byref(d)), "bvh shape")
    return n.value, d.value


# core/bvh.py:2: BVHNode's default leaf size (the only value below the root,
# since the recursion does not forward the argument, :12-13)
OBJECT_MEDIAN_LEAF_SIZE = 4


def object_median_shape(n_tris: int, max_leaf: int = OBJECT_MEDIAN_LEAF_SIZE) -> tuple[int, int]:
    """(node count, depth) of core/bvh.py's tree over n_tris triangles (the
    split is always at len // 2, so the shape depends only on the count)."""
    n, d = C.c_int(), C.c_int()
    nat.check(nat.lib().rx_bvh_object_median_shape(int(n_tris), int(max_leaf), C.byref(n), C.byref(d)),
              "bvh shape")
    return n.value, d.value


def build_object_median(v0, v1, v2, max_leaf_size=OBJECT_MEDIAN_LEAF_SIZE, device="cuda"):
    """core/bvh.py's ``BVHNode(triangles, max_leaf_size=...)`` over one
    mesh's triangles, built on the GPU and flattened in preorder.

    v0, v1, v2: (n, 3) vertex arrays, taken as fp64 (the reference's
    Python floats; scene/mesh.py:68-79 bounds and centroids are computed in
    fp64).  Returns a ``GPUBVHBuilder`` holding the tree in the same
    layouts as the other builds (``reference_arrays()``, ``ordered_tris``,
    the packed nodes), so it can be traversed like any other."""
    v = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1, 3)) for a in (v0, v1, v2)]
    n = len(v[0])
    if n == 0:
        raise ValueError("core/bvh.py builds over at least one triangle")
    z = np.zeros((n, 3), np.float32)
    batch = TriangleBatch(v[0].astype(np.float32), v[1].astype(np.float32), v[2].astype(np.float32), z, z, z,
                          np.zeros(n, np.int32), z + 0.5, np.zeros(n, np.float32), np.ones(n, np.float32),
                          np.zeros(n, np.float32))
    b = GPUBVHBuilder(device=device)
    b.OBJECT_MEDIAN_LEAF_SIZE = int(max_leaf_size)
    b._reset_for_build(batch)
    b._build_object_median_gpu(v64=v)
    b._pack_triangles()
    return b


class GPUBVHBuilder:
    """Builds the BVH on the GPU and keeps it resident (bvh_builder.py:275).

    Triangles are reordered into BVH leaf order; materials are
    de-duplicated into a shared palette in first-appearance order.
    """

    MAX_LEAF_SIZE = 8
    MAX_DEPTH = 32
    BUILD_METHOD = "morton"
    MORTON_BITS = 10
    # method="object_median": core/bvh.py's split rule, root leaf size
    OBJECT_MEDIAN_LEAF_SIZE = OBJECT_MEDIAN_LEAF_SIZE

    def __init__(self, device="cuda"):
        self.device = device
        self.mat_palette = []
        self._mat_key_to_idx = {}
        self._batch = TriangleBatch.empty()
        self._source_mat_idx = np.zeros(0, dtype=np.int32)
        self._n_nodes = 0
        self.depth = 0
        self._nodes_cache = None
        self._order_cache = np.zeros(0, dtype=np.int32)
        self.d = {}  # device tensors
        self.timings = {}

    # ---------------------------------------------------------------- build
    def build(self, tri_data, method=None):
        """Build from TriangleBatch objects or legacy triangle dicts."""
        batch = TriangleBatch.concat(self._coerce_batches(tri_data))
        self._reset_for_build(batch)
        n = batch.triangle_count
        if n:
            method = method or self.BUILD_METHOD
            if method == "morton":
                self._build_morton_gpu()
            elif method == "median":
                self._build_median_host()
            elif method == "object_median":
                self._build_object_median_gpu()
            else:
                raise ValueError(f"Unknown BVH build method: {method}")
            self._pack_triangles()
        return self

    def _coerce_batches(self, tri_data):
        if isinstance(tri_data, TriangleBatch):
            return [tri_data]
        if not tri_data:
            return []
        if all(isinstance(item, TriangleBatch) for item in tri_data):
            return list(tri_data)
        return [TriangleBatch.from_dicts(tri_data)]

    def _reset_for_build(self, batch):
        self.mat_palette = []
        self._mat_key_to_idx = {}
        self._batch = batch
        self._n_nodes = 0
        self.depth = 0
        self._nodes_cache = None
        self._order_cache = None
        self.d = {}
        self._source_mat_idx = self._build_material_indices(batch)

    def _build_material_indices(self, batch):
        if batch.triangle_count == 0:
            return np.zeros(0, dtype=np.int32)
        if batch.mat_idx is not None:
            remap = np.asarray([self._get_or_add_material(m) for m in batch.mat_palette], dtype=np.int32)
            return remap[batch.mat_idx] if len(remap) else np.zeros(0, dtype=np.int32)
        # per-triangle materials: de-duplicate whole rows at numpy speed, then
        # register the unique rows in first-appearance order
        rows = np.concatenate([batch.mat_type[:, None].view(np.float32), batch.albedo, batch.roughness[:, None],
                               batch.ior[:, None], batch.emission[:, None],
                               batch.texture_idx[:, None].view(np.float32)], axis=1)
        _, first, inverse = np.unique(rows.view(np.uint32), axis=0, return_index=True, return_inverse=True)
        appearance = np.argsort(first, kind="stable")
        local = np.empty(len(first), dtype=np.int32)
        for rank, u in enumerate(appearance):
            i = int(first[u])
            local[u] = self._get_or_add_material({
                "mat_type": int(batch.mat_type[i]), "albedo": batch.albedo[i].tolist(),
                "roughness": float(batch.roughness[i]), "ior": float(batch.ior[i]),
                "emission": float(batch.emission[i]), "texture_idx": int(batch.texture_idx[i])})
        return local[inverse.reshape(-1)]

    def _get_or_add_material(self, mat):
        key = material_key(mat)
        if key not in self._mat_key_to_idx:
            self._mat_key_to_idx[key] = len(self.mat_palette)
            self.mat_palette.append({"type": key[0], "albedo": list(key[1]), "roughness": key[2], "ior": key[3],
                                     "emission": key[4], "texture_idx": key[5]})
        return self._mat_key_to_idx[key]

    def _upload_geometry(self):
        import torch

        b = self._batch
        dev = self.device
        self.d["v0"] = torch.from_numpy(np.ascontiguousarray(b.v0)).to(dev)
        self.d["v1"] = torch.from_numpy(np.ascontiguousarray(b.v1)).to(dev)
        self.d["v2"] = torch.from_numpy(np.ascontiguousarray(b.v2)).to(dev)

    def _alloc_ref(self, n_nodes):
        import torch

        dev = self.device
        self.d["aabb_min"] = torch.empty((n_nodes, 3), dtype=torch.float32, device=dev)
        self.d["aabb_max"] = torch.empty((n_nodes, 3), dtype=torch.float32, device=dev)
        for k in ("left", "right", "tri_start", "tri_count"):
            self.d[k] = torch.empty(n_nodes, dtype=torch.int32, device=dev)
        self.d["nodes"] = torch.empty(n_nodes * nat.NODE_BYTES, dtype=torch.uint8, device=dev)
        self.d["header"] = torch.empty(nat.HEADER_BYTES, dtype=torch.uint8, device=dev)

    def _ref_struct(self):
        d = self.d
        return nat.RxBvhRef(d["aabb_min"].data_ptr(), d["aabb_max"].data_ptr(), d["left"].data_ptr(),
                            d["right"].data_ptr(), d["tri_start"].data_ptr(), d["tri_count"].data_ptr())

    def _build_morton_gpu(self):
        """bvh_builder.py:412-486 as kernels (rx_bvh_build_morton)."""
        import torch

        n = self._batch.triangle_count
        leaf = int(self.MAX_LEAF_SIZE)
        self._n_nodes, self.depth = morton_shape(n, leaf)
        self._upload_geometry()
        self._alloc_ref(self._n_nodes)
        ws = C.c_size_t()
        nat.check(nat.lib().rx_bvh_workspace_size(n, leaf, C.byref(ws)), "bvh workspace")
        work = torch.empty(ws.value, dtype=torch.uint8, device=self.device)
        self.d["order"] = torch.empty(n, dtype=torch.int32, device=self.device)
        ref = self._ref_struct()
        nat.check(nat.lib().rx_bvh_build_morton(
            self.d["v0"].data_ptr(), self.d["v1"].data_ptr(), self.d["v2"].data_ptr(), n, leaf,
            work.data_ptr(), ws.value, C.byref(ref), self.d["nodes"].data_ptr(), self.d["header"].data_ptr(),
            self.d["order"].data_ptr(), nat.stream_ptr()), "bvh build")
        self._workspace = work  # kept until the stream has consumed it

    def _build_object_median_gpu(self, v64=None):
        """core/bvh.py:2-48 as kernels (rx_bvh_build_object_median), over
        the batch's vertices widened to fp64 (exact), or over ``v64`` (three
        (n, 3) fp64 arrays) when the caller holds the fp64 originals."""
        import torch

        n = self._batch.triangle_count
        leaf = int(self.OBJECT_MEDIAN_LEAF_SIZE)
        self._n_nodes, self.depth = object_median_shape(n, leaf)
        self._upload_geometry()
        self._alloc_ref(self._n_nodes)
        if v64 is None:
            verts = [self.d[k].double().contiguous() for k in ("v0", "v1", "v2")]
        else:
            verts = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)).to(self.device)
                     for a in v64]
        ws = C.c_size_t()
        nat.check(nat.lib().rx_bvh_object_median_workspace_size(n, leaf, C.byref(ws)), "bvh workspace")
        work = torch.empty(ws.value, dtype=torch.uint8, device=self.device)
        self.d["order"] = torch.empty(n, dtype=torch.int32, device=self.device)
        ref = self._ref_struct()
        # returns once the build has finished on the stream
        nat.check(nat.lib().rx_bvh_build_object_median(
            verts[0].data_ptr(), verts[1].data_ptr(), verts[2].data_ptr(), n, leaf, work.data_ptr(), ws.value,
            C.byref(ref), self.d["nodes"].data_ptr(), self.d["header"].data_ptr(), self.d["order"].data_ptr(),
            nat.stream_ptr()), "bvh build")

    def _build_median_host(self):
        """bvh_builder.py:488-525: host recursion with numpy argpartition,
        then the nodes are packed on the device."""
        import torch

        b = self._batch
        stacked = np.stack([b.v0, b.v1, b.v2], axis=1)
        eps = np.float32(1e-4)
        lo = stacked.min(axis=1) - eps
        hi = stacked.max(axis=1) + eps
        cen = (b.v0 + b.v1 + b.v2) / np.float32(3.0)
        leaf, max_depth = int(self.MAX_LEAF_SIZE), int(self.MAX_DEPTH)
        amin, amax, left, right, start, count, chunks = [], [], [], [], [], [], []
        placed = 0
        deepest = 0
        work = [(np.arange(b.triangle_count, dtype=np.int32), 0, -1, 0)]
        while work:
            idx, depth, parent, side = work.pop()
            me = len(left)
            deepest = max(deepest, depth)
            amin.append(lo[idx].min(axis=0))
            amax.append(hi[idx].max(axis=0))
            left.append(-1)
            right.append(-1)
            start.append(-1)
            count.append(0)
            if parent >= 0:
                (left if side == 0 else right)[parent] = me
            if len(idx) <= leaf or depth >= max_depth:
                start[me] = placed
                count[me] = len(idx)
                placed += len(idx)
                chunks.append(idx)
                continue
            ext = (amax[me].astype(np.float64) - amin[me].astype(np.float64)).tolist()
            axis = 0 if (ext[0] >= ext[1] and ext[0] >= ext[2]) else (1 if ext[1] >= ext[2] else 2)
            mid = len(idx) // 2
            part = idx[np.argpartition(cen[idx, axis], mid)]
            work.append((part[mid:], depth + 1, me, 1))
            work.append((part[:mid], depth + 1, me, 0))
        if max(count) > nat.RX_MAX_LEAF_TRIS:
            raise AssertionError(f"BVH leaf of {max(count)} triangles exceeds {nat.RX_MAX_LEAF_TRIS}")
        self._n_nodes, self.depth = len(left), deepest
        self._alloc_ref(self._n_nodes)
        dev = self.device
        self.d["aabb_min"].copy_(torch.from_numpy(np.asarray(amin, np.float32)))
        self.d["aabb_max"].copy_(torch.from_numpy(np.asarray(amax, np.float32)))
        for k, v in (("left", left), ("right", right), ("tri_start", start), ("tri_count", count)):
            self.d[k].copy_(torch.from_numpy(np.asarray(v, np.int32)))
        self.d["order"] = torch.from_numpy(np.concatenate(chunks).astype(np.int32)).to(dev)
        self._upload_geometry()
        ref = self._ref_struct()
        nat.check(nat.lib().rx_bvh_pack(C.byref(ref), self._n_nodes, self.depth, self.d["nodes"].data_ptr(),
                                        self.d["header"].data_ptr(), nat.stream_ptr()), "bvh pack")

    def _pack_triangles(self):
        """Leaf-order gather + edge precompute (bvh_builder.py:527-544)."""
        import torch

        b = self._batch
        n = b.triangle_count
        dev = self.device
        d = self.d
        d["n0"] = torch.from_numpy(np.ascontiguousarray(b.n0)).to(dev)
        d["n1"] = torch.from_numpy(np.ascontiguousarray(b.n1)).to(dev)
        d["n2"] = torch.from_numpy(np.ascontiguousarray(b.n2)).to(dev)
        d["mat_idx"] = torch.from_numpy(np.ascontiguousarray(self._source_mat_idx, dtype=np.int32)).to(dev)
        d["has_uv"] = torch.from_numpy(np.ascontiguousarray(b.has_uv, dtype=np.int32)).to(dev)
        d["tris"] = torch.empty(n * nat.TRI_BYTES, dtype=torch.uint8, device=dev)
        d["shade"] = torch.empty(n * nat.SHADE_BYTES, dtype=torch.uint8, device=dev)
        nat.check(nat.lib().rx_bvh_pack_triangles(
            d["order"].data_ptr(), n, d["v0"].data_ptr(), d["v1"].data_ptr(), d["v2"].data_ptr(),
            d["n0"].data_ptr(), d["n1"].data_ptr(), d["n2"].data_ptr(), d["mat_idx"].data_ptr(),
            d["has_uv"].data_ptr(), d["tris"].data_ptr(), d["shade"].data_ptr(), nat.stream_ptr()), "pack tris")
        pal = np.zeros(max(len(self.mat_palette), 1), dtype=MATERIAL_DTYPE)
        for i, m in enumerate(self.mat_palette):
            pal[i] = (m["type"], m["albedo"], m["roughness"], m["ior"], m["emission"], m["texture_idx"])
        d["materials"] = torch.from_numpy(pal.view(np.uint8)).to(dev)

    def pack_uvs(self):
        """Per-corner UVs in leaf order, 6 floats per triangle (the
        _bvh_uv0..2 fields of fields.py:65-67), for textured materials."""
        import torch

        b = self._batch
        n = b.triangle_count
        if n == 0:
            return None
        dev = self.device
        uv = [torch.from_numpy(np.ascontiguousarray(getattr(b, k), dtype=np.float32)).to(dev)
              for k in ("uv0", "uv1", "uv2")]
        self.d["tri_uv"] = torch.empty(6 * n, dtype=torch.float32, device=dev)
        nat.check(nat.lib().rx_bvh_pack_uvs(self.d["order"].data_ptr(), n, uv[0].data_ptr(), uv[1].data_ptr(),
                                            uv[2].data_ptr(), self.d["tri_uv"].data_ptr(), nat.stream_ptr()),
                  "pack uvs")
        return self.d["tri_uv"]

    # ---------------------------------------------------------------- upload
    def upload(self):
        """The tree is already resident; check the reference's capacity
        limits (bvh_builder.py:575-580)."""
        assert self._n_nodes <= MAX_BVH_NODES, f"BVH has {self._n_nodes} nodes, limit is {MAX_BVH_NODES}"
        assert self.triangle_count <= MAX_TRIANGLES, \
            f"BVH has {self.triangle_count} triangles, limit is {MAX_TRIANGLES}"
        assert len(self.mat_palette) <= MAX_MATERIALS, \
            f"BVH has {len(self.mat_palette)} materials, limit is {MAX_MATERIALS}"
        return self

    # ---------------------------------------------------------------- results
    @property
    def triangle_count(self):
        return self._batch.triangle_count

    @property
    def node_count(self):
        return self._n_nodes

    @property
    def ordered_tris(self):
        if self._order_cache is None:
            self._order_cache = (self.d["order"].cpu().numpy() if "order" in self.d
                                 else np.zeros(0, dtype=np.int32))
        return self._order_cache

    def reference_arrays(self):
        """Node arrays in the reference layout (fields.py:50-56), read back."""
        if not self._n_nodes:
            return {k: np.zeros((0, 3) if k.startswith("aabb") else 0, np.float32 if k.startswith("aabb")
                                else np.int32) for k in ("aabb_min", "aabb_max", "left", "right", "tri_start",
                                                         "tri_count")}
        return {k: self.d[k].cpu().numpy() for k in ("aabb_min", "aabb_max", "left", "right", "tri_start",
                                                     "tri_count")}

    @property
    def nodes(self):
        """Node dicts as the reference's ``nodes`` list (read back lazily)."""
        if self._nodes_cache is None:
            r = self.reference_arrays()
            self._nodes_cache = [
                {"aabb_min": r["aabb_min"][i].tolist(), "aabb_max": r["aabb_max"][i].tolist(),
                 "left": int(r["left"][i]), "right": int(r["right"][i]), "tri_start": int(r["tri_start"][i]),
                 "tri_count": int(r["tri_count"][i])}
                for i in range(self._n_nodes)]
        return self._nodes_cache

    def _ordered(self, name):
        return getattr(self._batch, name)[self.ordered_tris]

    _ordered_v0 = property(lambda self: self._ordered("v0"))
    _ordered_v1 = property(lambda self: self._ordered("v1"))
    _ordered_v2 = property(lambda self: self._ordered("v2"))
    _ordered_n0 = property(lambda self: self._ordered("n0"))
    _ordered_n1 = property(lambda self: self._ordered("n1"))
    _ordered_n2 = property(lambda self: self._ordered("n2"))
    _ordered_uv0 = property(lambda self: self._ordered("uv0"))
    _ordered_uv1 = property(lambda self: self._ordered("uv1"))
    _ordered_uv2 = property(lambda self: self._ordered("uv2"))
    _ordered_has_uv = property(lambda self: self._ordered("has_uv"))
    _ordered_mat_idx = property(lambda self: self._source_mat_idx[self.ordered_tris])

    def _longest_axis(self, aabb_min, aabb_max):
        ex, ey, ez = np.asarray(aabb_max) - np.asarray(aabb_min)
        if ex >= ey and ex >= ez:
            return 0
        if ey >= ez:
            return 1
        return 2
