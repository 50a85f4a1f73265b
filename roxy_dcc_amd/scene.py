"""Device-resident scene: the flattened data render_kernel reads.

``FlatScene`` is the host-side result of scene flattening
(rendering/taichi/extractor.py:84-249): triangle batches for the BVH,
primitive slots (fields.py:22-40) and emissive-sphere lights
(fields.py:43-47, extractor.py:36-50).  ``DeviceScene`` uploads it once:
the BVH is built on the GPU, primitives and lights are packed into
48-/32-byte records, and ``struct()`` returns the ``rx_scene`` view the
kernels take.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _native as nat
from .bvh import MATERIAL_DTYPE, GPUBVHBuilder, TriangleBatch

PRIM_DTYPE = np.dtype([("a", "<f4", 4), ("b", "<f4", 4), ("c", "<f4", 3), ("type", "<i4")])
LIGHT_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("albedo", "<f4", 3), ("emission", "<f4")])
TEXTURE_DTYPE = np.dtype([("width", "<i4"), ("height", "<i4"), ("flip_v", "<i4"), ("pad0", "<i4"),
                          ("offset", "<i8"), ("pad1", "<i8")])
assert PRIM_DTYPE.itemsize == nat.PRIM_BYTES and LIGHT_DTYPE.itemsize == nat.LIGHT_BYTES

# capacity limits of the reference fields (rendering/taichi/fields.py:3, :11)
MAX_OBJECTS = 10_000_000
MAX_LIGHTS = 4_096
MAX_TEXTURE_SIZE = 2048        # rendering/taichi/fields.py: textures are downsampled to fit


@dataclass
class PrimitiveSlots:
    """Analytic primitive slots in the reference field layout (fields.py:22-40)."""

    obj_type: np.ndarray
    center: np.ndarray
    radius: np.ndarray
    normal: np.ndarray
    offset: np.ndarray
    extra: np.ndarray
    v0: np.ndarray
    v1: np.ndarray
    v2: np.ndarray
    mat_type: np.ndarray
    albedo: np.ndarray
    roughness: np.ndarray
    ior: np.ndarray
    emission: np.ndarray

    @classmethod
    def from_dicts(cls, slots):
        """From the dicts SceneObject.taichi_export produces
        (scene/scene_object.py:317-410)."""
        z3 = [0.0, 0.0, 0.0]

        def col(k, default=None, dt=np.float32, shape=None):
            a = np.asarray([s.get(k, default) if default is not None else s[k] for s in slots], dtype=dt)
            return a.reshape(shape) if shape else a

        if not slots:
            return cls.empty()
        return cls(col("type", dt=np.int32), col("center", shape=(-1, 3)), col("radius"),
                   col("normal", shape=(-1, 3)), col("offset"), col("extra", shape=(-1, 3)),
                   col("v0", z3, shape=(-1, 3)), col("v1", z3, shape=(-1, 3)), col("v2", z3, shape=(-1, 3)),
                   col("mat_type", dt=np.int32), col("albedo", shape=(-1, 3)), col("roughness"), col("ior"),
                   col("emission"))

    @classmethod
    def empty(cls):
        z3 = np.zeros((0, 3), np.float32)
        z = np.zeros(0, np.float32)
        zi = np.zeros(0, np.int32)
        return cls(zi, z3, z, z3, z, z3, z3, z3, z3, zi, z3, z, z, z)

    def __len__(self):
        return len(self.obj_type)

    def lights(self):
        """Emissive spheres with positive emission (extractor.py:36-40)."""
        sel = (self.obj_type == 0) & (self.mat_type == 3) & (self.emission > 0.0)
        return Lights(self.center[sel], self.radius[sel], self.albedo[sel], self.emission[sel])

    def records(self):
        out = np.zeros(len(self), dtype=PRIM_DTYPE)
        out["type"] = self.obj_type
        for i, t in enumerate(self.obj_type):
            if t == nat.RX_PRIM_SPHERE:
                out["a"][i] = [*self.center[i], self.radius[i]]
            elif t == nat.RX_PRIM_PLANE:
                out["a"][i] = [*self.normal[i], self.offset[i]]
            elif t == nat.RX_PRIM_CUBE:
                out["a"][i] = [*self.center[i], 0.0]
                out["b"][i] = [*self.extra[i], 0.0]
            elif t == nat.RX_PRIM_TRIANGLE:
                v = np.concatenate([self.v0[i], self.v1[i], self.v2[i]])
                out["a"][i] = v[0:4]
                out["b"][i] = v[4:8]
                out["c"][i] = [v[8], 0.0, 0.0]
        return out

    def material_records(self):
        out = np.zeros(len(self), dtype=MATERIAL_DTYPE)
        out["type"] = self.mat_type
        out["albedo"] = self.albedo
        out["roughness"] = self.roughness
        out["ior"] = self.ior
        out["emission"] = self.emission
        out["texture"] = -1
        return out


@dataclass
class Lights:
    center: np.ndarray
    radius: np.ndarray
    albedo: np.ndarray
    emission: np.ndarray

    def __len__(self):
        return len(self.radius)

    def records(self):
        out = np.zeros(len(self), dtype=LIGHT_DTYPE)
        out["center"] = self.center
        out["radius"] = self.radius
        out["albedo"] = self.albedo
        out["emission"] = self.emission
        return out


@dataclass
class FlatScene:
    """Flattened world-space scene (the output of scene extraction)."""

    batches: list = field(default_factory=list)
    prims: PrimitiveSlots = field(default_factory=PrimitiveSlots.empty)
    textures: list = field(default_factory=list)     # ImageTexture, indexed by material texture_idx

    @property
    def triangle_count(self):
        return sum(b.triangle_count for b in self.batches)

    def triangles(self) -> TriangleBatch:
        return TriangleBatch.concat(self.batches)


class DeviceScene:
    """A FlatScene resident in HBM, ready for the ray kernels."""

    def __init__(self, flat: FlatScene, device="cuda", max_leaf=None, method=None):
        import torch

        nat.require_gpu()
        self.flat = flat
        self.device = device
        n_prims = len(flat.prims)
        assert n_prims <= MAX_OBJECTS, f"Scene has {n_prims} primitives but MAX_OBJECTS={MAX_OBJECTS}"
        lights = flat.prims.lights()
        assert len(lights) <= MAX_LIGHTS, f"Scene has {len(lights)} emissive sphere lights but MAX_LIGHTS={MAX_LIGHTS}"
        self.lights = lights
        self.builder = GPUBVHBuilder(device=device)
        if max_leaf is not None:
            self.builder.MAX_LEAF_SIZE = int(max_leaf)
        self.builder.build(flat.batches, method=method)
        self.builder.upload()
        # finite vertices make finite boxes: rx_render may then trace a copy
        # of the tree without child links (RX_SCENE_FINITE_BOXES)
        self.finite = all(bool(np.isfinite(getattr(b, k)).all()) for b in flat.batches for k in ("v0", "v1", "v2"))

        def dev(arr):
            if len(arr) == 0:
                return torch.zeros(64, dtype=torch.uint8, device=device)
            return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).reshape(-1)).to(device)

        self.d_prims = dev(flat.prims.records())
        self.d_prim_mats = dev(flat.prims.material_records())
        self.d_lights = dev(lights.records())
        # texture cache (extractor.py:53-81): one u8 RGB pool, a 32-byte
        # record per texture, and leaf-order UVs for the triangles
        self.n_textures = len(flat.textures)
        self.d_textures = self.d_texels = self.d_tri_uv = None
        if self.n_textures:
            recs = np.zeros(self.n_textures, dtype=TEXTURE_DTYPE)
            chunks, off = [], 0
            for i, tex in enumerate(flat.textures):
                px = np.ascontiguousarray(tex.load_pixels_u8(max_size=MAX_TEXTURE_SIZE), dtype=np.uint8)
                h, w = px.shape[:2]
                recs[i] = (w, h, 1 if tex.flip_v else 0, 0, off, 0)
                chunks.append(px.reshape(-1))
                off += px.size
            self.d_texels = torch.from_numpy(np.concatenate(chunks)).to(device)
            self.d_textures = dev(recs)
            if self.builder.triangle_count:
                self.d_tri_uv = self.builder.pack_uvs()
        self._struct = None

    @property
    def n_prims(self):
        return len(self.flat.prims)

    @property
    def n_lights(self):
        return len(self.lights)

    @property
    def n_tris(self):
        return self.builder.triangle_count

    def struct(self) -> nat.RxScene:
        if self._struct is None:
            b = self.builder
            has_bvh = b.triangle_count > 0
            d = b.d
            self._struct = nat.RxScene(
                d["nodes"].data_ptr() if has_bvh else None,
                d["header"].data_ptr() if has_bvh else None,
                d["tris"].data_ptr() if has_bvh else None,
                d["shade"].data_ptr() if has_bvh else None,
                d["materials"].data_ptr() if has_bvh else None,
                self.d_prims.data_ptr(), self.d_prim_mats.data_ptr(), self.d_lights.data_ptr(),
                self.d_tri_uv.data_ptr() if self.d_tri_uv is not None else None,
                self.d_textures.data_ptr() if self.d_textures is not None else None,
                self.d_texels.data_ptr() if self.d_texels is not None else None,
                b.node_count, b.triangle_count, len(b.mat_palette), self.n_prims, self.n_lights,
                b.depth, self.n_textures, nat.RX_SCENE_FINITE_BOXES if self.finite else 0)
        return self._struct

    def stats(self):
        b = self.builder
        return {
            "primitive_count": self.n_prims,
            "bvh_nodes": b.node_count,
            "bvh_triangles": b.triangle_count,
            "bvh_materials": len(b.mat_palette) if b.triangle_count else 0,
            "bvh_leaf_size": b.MAX_LEAF_SIZE if b.triangle_count else 0,
        }

    def hbm_bytes(self) -> int:
        """Device bytes this scene keeps resident."""
        total = sum(t.numel() * t.element_size() for t in self.builder.d.values())
        extra = [t for t in (self.d_prims, self.d_prim_mats, self.d_lights, self.d_textures, self.d_texels)
                 if t is not None]
        return total + sum(t.numel() * t.element_size() for t in extra)
