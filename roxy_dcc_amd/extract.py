"""Scene flattening: rendering/taichi/extractor.py:84-249.

Walks the object tree exactly as ``_collect`` does (:134-205): a ``Mesh``
(per-triangle objects) is transformed triangle by triangle, an
``IndexedMesh`` by the library's native flattener (``rx_flatten_indexed``,
the fp64 transform of ``indexed_triangle_arrays`` on every host core), and any
other geometry is exported once per object with ``taichi_export()``.
Lights are the exported spheres with material type 3 and emission > 0
(:36-40).  Per-triangle ``Mesh`` data is appended as one batch after the
indexed batches (:223-224), as in the reference.

``extract_scene`` returns the reference's statistics dict
(:232-249); the device-resident scene it built rides along as the
``scene`` attribute of that dict.
"""
from __future__ import annotations

import os

import numpy as np

from . import _native as nat
from .bvh import TriangleBatch
from .scene import DeviceScene, FlatScene, PrimitiveSlots


class ExtractResult(dict):
    """The statistics dict of extractor.py:232-249, plus ``.scene``."""

    scene: DeviceScene = None
    flat: FlatScene = None


# Texture cache capacity (rendering/taichi/fields.py:9-10).
MAX_TEXTURES = 8
MAX_TEXTURE_SIZE = 2048


class _Textures:
    """The texture list of extractor.py:93-117: one slot per distinct
    (path, flip_v, identity of an in-memory texture), first-use order."""

    def __init__(self):
        self.items = []
        self.index = {}

    def idx(self, mat):
        texture = getattr(mat, "_albedo_texture", None)
        if texture is None:
            return -1
        key = (texture.path, bool(texture.flip_v), id(texture) if texture.path is None else None)
        if key in self.index:
            return self.index[key]
        if len(self.items) >= MAX_TEXTURES:
            raise ValueError(f"Scene has more than {MAX_TEXTURES} image textures; increase MAX_TEXTURES")
        self.index[key] = len(self.items)
        self.items.append(texture)
        return self.index[key]


def _material_data(mat, textures):
    mt = mat.taichi_type_id()
    params = mat.taichi_params()
    return {
        "mat_type": mt,
        "albedo": [float(c) for c in mat._albedo],
        "roughness": params[0] if mt in (1, 4) else 0.0,
        "ior": params[0] if mt == 2 else 1.0,
        "emission": params[0] if mt == 3 else 0.0,
        "texture_idx": textures.idx(mat),
    }


def _default_material():
    from .world import Diffuse

    return Diffuse((0.8, 0.8, 0.8))


def _is_indexed(geo):
    return hasattr(geo, "_tri_pos_idx") and hasattr(geo, "indexed_triangle_arrays")


def _is_object_mesh(geo):
    return hasattr(geo, "_triangles")


def _host_threads():
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return n if n > 0 else min(16, os.cpu_count() or 1)


def _ptr(a):
    return None if a is None else a.ctypes.data


def indexed_arrays_native(geo, matrix=None, normal_matrix=None):
    """``IndexedMesh.indexed_triangle_arrays`` (scene/mesh.py:397-430) by
    the library's native flattener: same keys, fp32 arrays."""
    n = geo.triangle_count
    out = {k: np.empty((n, 3), np.float32) for k in ("v0", "v1", "v2", "n0", "n1", "n2")}
    out.update({k: np.empty((n, 2), np.float32) for k in ("uv0", "uv1", "uv2")})
    has_uv = np.empty(n, np.int32)
    pos = np.ascontiguousarray(geo._positions, dtype=np.float64)
    tpi = np.ascontiguousarray(geo._tri_pos_idx, dtype=np.int32)
    nrm = None if geo._normals is None else np.ascontiguousarray(geo._normals, dtype=np.float64)
    tni = np.ascontiguousarray(geo._tri_normal_idx, dtype=np.int32)
    uvs = None if geo._uvs is None else np.ascontiguousarray(geo._uvs, dtype=np.float64)
    tui = np.ascontiguousarray(geo._tri_uv_idx, dtype=np.int32)
    m = None if matrix is None else np.ascontiguousarray(matrix.rows, dtype=np.float64)
    nm = None if normal_matrix is None else np.ascontiguousarray(normal_matrix.rows, dtype=np.float64)
    nat.check(nat.lib().rx_flatten_indexed(
        _ptr(pos), len(pos), _ptr(tpi), n, _ptr(nrm), 0 if nrm is None else len(nrm), _ptr(tni),
        _ptr(uvs), 0 if uvs is None else len(uvs), _ptr(tui), _ptr(m), _ptr(nm),
        *[_ptr(out[k]) for k in ("v0", "v1", "v2", "n0", "n1", "n2", "uv0", "uv1", "uv2")], _ptr(has_uv),
        _host_threads()), "flatten")
    out["has_uv"] = has_uv
    out["group_idx"] = np.ascontiguousarray(geo._tri_group_idx, dtype=np.int32)
    out["groups"] = list(geo._groups)
    return out


def _uv2(uv):
    return [float(uv[0]), float(uv[1])] if uv is not None else [0.0, 0.0]


def flatten_scene(world) -> FlatScene:
    """Host-side flattening of ``world`` (no device work)."""
    prim_slots = []
    tri_dicts = []
    batches = []
    textures = _Textures()

    def collect(obj):
        if not obj.renderable:
            return
        for shape in obj.shapes:
            geo = shape.geometry
            if _is_object_mesh(geo):
                m = obj.world_matrix
                inv_t = obj.world_inverse_transpose_matrix
                for tri in geo._triangles:
                    def xn(n):
                        v = np.asarray(list(inv_t.transform_vector(n)), dtype=np.float64)
                        ln = float(np.sqrt(v @ v))
                        return list(v / ln) if ln > 0 else [0.0, 0.0, 0.0]

                    face = tri._normal
                    mat = shape.material_for_group(tri._group) or _default_material()
                    tri_dicts.append({
                        "v0": list(m.transform_point(tri._v0)), "v1": list(m.transform_point(tri._v1)),
                        "v2": list(m.transform_point(tri._v2)),
                        "n0": xn(tri._n0 if tri._n0 is not None else face),
                        "n1": xn(tri._n1 if tri._n1 is not None else face),
                        "n2": xn(tri._n2 if tri._n2 is not None else face),
                        "uv0": _uv2(tri._uv0), "uv1": _uv2(tri._uv1), "uv2": _uv2(tri._uv2),
                        "has_uv": int(tri._uv0 is not None and tri._uv1 is not None and tri._uv2 is not None),
                        **_material_data(mat, textures),
                    })
            elif _is_indexed(geo):
                arrays = indexed_arrays_native(geo, obj.world_matrix, obj.world_inverse_transpose_matrix)
                by_group = {}
                for gi, group in enumerate(arrays["groups"]):
                    mat = shape.material_for_group(group) or _default_material()
                    by_group[gi] = _material_data(mat, textures)
                batches.append(TriangleBatch.from_indexed_arrays(arrays, by_group))
            else:
                exported = obj.taichi_export()
                if isinstance(exported, list):
                    prim_slots.extend(exported)
                elif exported:
                    prim_slots.append(exported)
                break  # taichi_export covers every shape of the object
        for child in obj.children:
            collect(child)

    for obj in world.objects:
        collect(obj)
    if tri_dicts:
        batches.append(TriangleBatch.from_dicts(tri_dicts))
    return FlatScene(batches=batches, prims=PrimitiveSlots.from_dicts(prim_slots), textures=textures.items)


def extract_scene(world, device="cuda", max_leaf=None, method=None) -> ExtractResult:
    """Flatten ``world``, build its BVH on the GPU and keep it resident."""
    flat = flatten_scene(world)
    scene = DeviceScene(flat, device=device, max_leaf=max_leaf, method=method)
    out = ExtractResult(scene.stats())
    out.scene = scene
    out.flat = flat
    return out
