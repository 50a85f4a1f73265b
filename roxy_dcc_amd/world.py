"""Minimal scene model with the attributes scene flattening reads.

The reference's scene graph (scene/world.py, scene/scene_object.py,
scene/shape.py, scene/camera.py, scene/materials.py, scene/primitives.py,
scene/mesh.py) cannot be imported without Taichi.  These classes carry
the same public attributes the render path reads, so ``extract_scene``
and ``TaichiRenderer`` accept either these objects or the reference's own
(duck-typed):

  World          .objects, .active_camera, ._use_sky, ._background_color
  SceneObject    .renderable, .shapes, .children, .world_matrix,
                 .world_inverse_matrix, .world_inverse_transpose_matrix,
                 .material, .taichi_export()
  Shape          .geometry, .material, .material_for_group(group)
  Camera         .position/.forward/.right/.up/.fov/.aspect_ratio with the
                 basis rule of scene/camera.py:32-34
  materials      taichi_type_id()/taichi_params() ids and parameters of
                 scene/materials.py:50-184 (roughness clamped to [0, 1])
  Sphere/Plane/Cube/IndexedMesh  geometry records
"""
from __future__ import annotations

import math

import numpy as np


# ------------------------------------------------------------------ math
class Mat4:
    """Row-major 4x4 (``rows`` like core/matrix.py) acting on column vectors."""

    def __init__(self, rows=None):
        self.rows = np.eye(4) if rows is None else np.asarray(rows, dtype=np.float64).reshape(4, 4)

    def __matmul__(self, other):
        return Mat4(self.rows @ other.rows)

    def transform_point(self, p):
        v = self.rows @ np.array([p[0], p[1], p[2], 1.0])
        if abs(v[3]) > 1e-12:
            v = v / v[3]
        return v[:3]

    def transform_vector(self, d):
        return self.rows[:3, :3] @ np.asarray(d, dtype=np.float64)

    def inverse(self):
        return Mat4(np.linalg.inv(self.rows))

    def transpose(self):
        return Mat4(self.rows.T)

    @staticmethod
    def trs(translation=(0, 0, 0), rotation_deg=(0, 0, 0), scale=(1, 1, 1)):
        """T * Rz * Ry * Rx * S (rotation in degrees about x, then y, then z)."""
        rx, ry, rz = (math.radians(a) for a in rotation_deg)
        cx, sx, cy, sy, cz, sz = math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry), math.cos(rz), math.sin(rz)
        r_x = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
        r_y = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
        r_z = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
        m = np.eye(4)
        m[:3, :3] = r_z @ r_y @ r_x @ np.diag(np.asarray(scale, dtype=np.float64))
        m[:3, 3] = translation
        return Mat4(m)


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    n = float(np.sqrt(v @ v))
    return v / n if n > 0 else np.zeros(3)


# ------------------------------------------------------------------ textures
class ImageTexture:
    """RGB image texture, repeating UVs, bilinear (scene/textures.py:6-70).
    Image files are decoded with pygame, as in the reference; an in-memory
    array needs nothing else."""

    def __init__(self, path=None, pixels=None, flip_v=True):
        if path is None and pixels is None:
            raise ValueError("ImageTexture needs either a path or pixel array")
        self.path = str(path) if path is not None else None
        self._pixels = None if pixels is None else self._coerce_pixels(pixels)
        self.flip_v = flip_v

    @classmethod
    def from_array(cls, pixels, flip_v=True):
        return cls(pixels=pixels, flip_v=flip_v)

    @staticmethod
    def _coerce_pixels(pixels):
        arr = np.asarray(pixels, dtype=np.float32)
        if arr.ndim != 3 or arr.shape[2] < 3:
            raise ValueError("texture pixels must be an HxWx3 array")
        arr = arr[:, :, :3]
        if arr.size and arr.max() > 1.0:
            arr = arr / 255.0
        return np.clip(arr, 0.0, 1.0)

    def _load_pixels(self):
        if self._pixels is not None:
            return self._pixels
        try:
            import pygame
        except ImportError as exc:
            raise RuntimeError("pygame is required to load image textures") from exc
        surface = pygame.image.load(self.path)
        pixels = pygame.surfarray.array3d(surface).astype(np.float32) / 255.0
        self._pixels = np.transpose(pixels[:, :, :3], (1, 0, 2))
        return self._pixels

    def load_pixels_u8(self, max_size=None):
        """uint8 HxWx3 for the texture cache (scene/textures.py:43-70)."""
        if self.path is None:
            return np.clip(self._load_pixels() * 255.0 + 0.5, 0, 255).astype(np.uint8)
        try:
            import pygame
        except ImportError as exc:
            raise RuntimeError("pygame is required to load image textures") from exc
        surface = pygame.image.load(self.path)
        width, height = surface.get_size()
        if max_size is not None and max(width, height) > max_size:
            scale = float(max_size) / float(max(width, height))
            surface = pygame.transform.smoothscale(
                surface, (max(1, int(round(width * scale))), max(1, int(round(height * scale)))))
        return np.transpose(pygame.surfarray.array3d(surface)[:, :, :3], (1, 0, 2)).copy()


# ------------------------------------------------------------------ materials
class Material:
    type_id = 0

    def __init__(self, albedo=(1.0, 1.0, 1.0), albedo_texture=None):
        self._albedo = tuple(float(c) for c in albedo)
        self._albedo_texture = albedo_texture

    def taichi_type_id(self):
        return self.type_id

    def taichi_params(self):
        return []


class Diffuse(Material):
    type_id = 0


class Metal(Material):
    type_id = 1

    def __init__(self, albedo, roughness=0.0, albedo_texture=None):
        super().__init__(albedo, albedo_texture)
        self._roughness = min(max(float(roughness), 0.0), 1.0)

    def taichi_params(self):
        return [self._roughness]


class Dielectric(Material):
    type_id = 2

    def __init__(self, ior=1.5, albedo=(1.0, 1.0, 1.0), albedo_texture=None):
        super().__init__(albedo, albedo_texture)
        self._ior = float(ior)

    def taichi_params(self):
        return [self._ior]


class Emissive(Material):
    type_id = 3

    def __init__(self, color, intensity=1.0, albedo_texture=None):
        super().__init__(color, albedo_texture)
        self._intensity = float(intensity)

    def taichi_params(self):
        return [self._intensity]


class Glossy(Material):
    type_id = 4

    def __init__(self, albedo, roughness=0.5, albedo_texture=None):
        super().__init__(albedo, albedo_texture)
        self._roughness = min(max(float(roughness), 0.0), 1.0)

    def taichi_params(self):
        return [self._roughness]


# ------------------------------------------------------------------ geometry
class Sphere:
    def __init__(self, radius=1.0):
        self._radius = float(radius)


class Plane:
    def __init__(self, normal=(0.0, 1.0, 0.0), distance=0.0):
        self._normal = _unit(normal)
        self._distance = float(distance)


class Cube:
    def __init__(self, side_length=1.0):
        self._side_length = float(side_length)


class IndexedMesh:
    """Shared vertex arrays plus per-triangle index tables (scene/mesh.py:263-296)."""

    def __init__(self, positions, tri_pos_idx, normals=None, tri_normal_idx=None, uvs=None, tri_uv_idx=None,
                 groups=None, tri_group_idx=None, name=""):
        self._positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        self._tri_pos_idx = np.asarray(tri_pos_idx, dtype=np.int32).reshape(-1, 3)
        t = len(self._tri_pos_idx)
        self._normals = None if normals is None else np.asarray(normals, dtype=np.float64).reshape(-1, 3)
        self._tri_normal_idx = (np.full((t, 3), -1, np.int32) if tri_normal_idx is None
                                else np.asarray(tri_normal_idx, np.int32).reshape(t, 3))
        self._uvs = None if uvs is None else np.asarray(uvs, dtype=np.float64).reshape(-1, 2)
        self._tri_uv_idx = (np.full((t, 3), -1, np.int32) if tri_uv_idx is None
                            else np.asarray(tri_uv_idx, np.int32).reshape(t, 3))
        self._groups = list(groups) if groups else ["default"]
        self._tri_group_idx = (np.zeros(t, np.int32) if tri_group_idx is None
                               else np.asarray(tri_group_idx, np.int32).reshape(-1))
        self._name = name

    @property
    def triangle_count(self):
        return len(self._tri_pos_idx)

    @property
    def vertex_count(self):
        return len(self._positions)

    @property
    def groups(self):
        return list(self._groups)

    def group_for_triangle(self, i):
        return self._groups[int(self._tri_group_idx[i])]

    def indexed_triangle_arrays(self, matrix=None, normal_matrix=None, dtype=np.float32):
        """World-space per-triangle arrays (scene/mesh.py:397-430): points
        through ``matrix`` with a perspective divide where |w| > 1e-12,
        normals through ``normal_matrix`` then normalised (rows with length
        <= 1e-12 left as they are); missing normals fall back to the face
        normal, missing UVs to zeros, all in fp64 before the final cast."""
        tv = self._positions[self._tri_pos_idx]
        tv = _xf_points(tv, matrix)
        tn = self._normal_table()
        tn = _xf_vectors(tn, normal_matrix)
        tn = _normalize_rows(tn.reshape(-1, 3)).reshape(-1, 3, 3)
        tuv = np.zeros((self.triangle_count, 3, 2))
        if self._uvs is not None:
            ok = np.all(self._tri_uv_idx >= 0, axis=1)
            if np.any(ok):
                tuv[ok] = self._uvs[self._tri_uv_idx[ok]]
        has_uv = np.all(self._tri_uv_idx >= 0, axis=1).astype(np.int32)
        return {
            "v0": tv[:, 0].astype(dtype), "v1": tv[:, 1].astype(dtype), "v2": tv[:, 2].astype(dtype),
            "n0": tn[:, 0].astype(dtype), "n1": tn[:, 1].astype(dtype), "n2": tn[:, 2].astype(dtype),
            "uv0": tuv[:, 0].astype(dtype), "uv1": tuv[:, 1].astype(dtype), "uv2": tuv[:, 2].astype(dtype),
            "has_uv": has_uv, "group_idx": self._tri_group_idx.astype(np.int32), "groups": list(self._groups),
        }

    def _normal_table(self):
        if self.triangle_count == 0:
            return np.zeros((0, 3, 3))
        tv = self._positions[self._tri_pos_idx]
        face = _normalize_rows(np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0]))
        tn = np.repeat(face[:, None, :], 3, axis=1)
        if self._normals is not None:
            ok = np.all(self._tri_normal_idx >= 0, axis=1)
            if np.any(ok):
                tn[ok] = self._normals[self._tri_normal_idx[ok]]
        return tn


def _rows(matrix):
    return None if matrix is None else np.asarray(matrix.rows, dtype=np.float64)


def _xf_points(pts, matrix):
    """scene/mesh.py:547-558 with the products spelled out elementwise in
    fp64: ((x m_j0 + y m_j1) + z m_j2) + m_j3, perspective divide where
    |w| > 1e-12.  (The reference writes a numpy matmul, whose rounding is the
    linked BLAS's; the native flattener uses exactly this order.)"""
    m = _rows(matrix)
    out = np.asarray(pts, dtype=np.float64)
    if m is None:
        return out.copy()
    flat = out.reshape(-1, 3)
    x, y, z = flat[:, 0], flat[:, 1], flat[:, 2]
    res = np.stack([((x * m[j, 0] + y * m[j, 1]) + z * m[j, 2]) + m[j, 3] for j in range(3)], axis=1)
    w = ((x * m[3, 0] + y * m[3, 1]) + z * m[3, 2]) + m[3, 3]
    div = np.abs(w) > 1e-12
    res[div] /= w[div, None]
    return res.reshape(out.shape)


def _xf_vectors(vecs, matrix):
    """scene/mesh.py:560-567, elementwise: (x m_j0 + y m_j1) + z m_j2."""
    m = _rows(matrix)
    out = np.asarray(vecs, dtype=np.float64)
    if m is None:
        return out.copy()
    flat = out.reshape(-1, 3)
    x, y, z = flat[:, 0], flat[:, 1], flat[:, 2]
    return np.stack([(x * m[j, 0] + y * m[j, 1]) + z * m[j, 2] for j in range(3)], axis=1).reshape(out.shape)


def _normalize_rows(rows):
    lengths = np.sqrt((rows[:, 0] * rows[:, 0] + rows[:, 1] * rows[:, 1]) + rows[:, 2] * rows[:, 2])
    out = rows.copy()
    ok = lengths > 1e-12
    out[ok] /= lengths[ok, None]
    return out


def indexed_triangle_arrays(mesh, matrix=None, normal_matrix=None, dtype=np.float32):
    return mesh.indexed_triangle_arrays(matrix, normal_matrix, dtype=dtype)


# ------------------------------------------------------------------ scene graph
class Shape:
    def __init__(self, geometry, material=None, group_materials=None):
        self._geometry = geometry
        self._material = material
        self._group_materials = dict(group_materials or {})

    @property
    def geometry(self):
        return self._geometry

    @property
    def material(self):
        return self._material

    def material_for_group(self, group_name):
        return self._group_materials.get(group_name, self._material)

    @property
    def material_groups(self):
        return dict(self._group_materials)


class SceneObject:
    def __init__(self, name="", shapes=None, transform=None, renderable=True):
        self._name = name
        self._shapes = list(shapes or [])
        self._local = transform if transform is not None else Mat4()
        self._children = []
        self._parent = None
        self._renderable = renderable

    def add_child(self, child):
        child._parent = self
        self._children.append(child)

    @property
    def name(self):
        return self._name

    @property
    def children(self):
        return list(self._children)

    @property
    def renderable(self):
        return self._renderable

    @property
    def shapes(self):
        return self._shapes

    @property
    def shape(self):
        return self._shapes[0].geometry if self._shapes else None

    @property
    def material(self):
        return self._shapes[0].material if self._shapes else None

    @property
    def world_matrix(self):
        return self._local if self._parent is None else self._parent.world_matrix @ self._local

    @property
    def world_inverse_matrix(self):
        return self.world_matrix.inverse()

    @property
    def world_inverse_transpose_matrix(self):
        return self.world_inverse_matrix.transpose()

    def taichi_export(self):
        """World-space primitive slot (scene/scene_object.py:317-410)."""
        shape = self.shape
        if shape is None:
            return []
        m = self.world_matrix
        mat = self.material or Diffuse((0.8, 0.8, 0.8))
        mt = mat.taichi_type_id()
        params = mat.taichi_params()
        md = {"albedo": list(mat._albedo), "mat_type": mt, "roughness": params[0] if mt in (1, 4) else 0.0,
              "ior": params[0] if mt == 2 else 1.0, "emission": params[0] if mt == 3 else 0.0}
        axis_len = [float(np.linalg.norm(m.transform_vector(e))) for e in np.eye(3)]
        if isinstance(shape, Sphere):
            return {"type": 0, "center": list(m.transform_point((0, 0, 0))), "radius": shape._radius * max(axis_len),
                    "normal": [0.0, 1.0, 0.0], "offset": 0.0, "extra": [0.0, 0.0, 0.0], **md}
        if isinstance(shape, Plane):
            n_world = _unit(self.world_inverse_matrix.transpose().transform_vector(shape._normal))
            p_world = m.transform_point(shape._normal * -shape._distance)
            return {"type": 1, "center": [0.0, 0.0, 0.0], "radius": 0.0, "normal": list(n_world),
                    "offset": float(-(n_world @ p_world)), "extra": [0.0, 0.0, 0.0], **md}
        if isinstance(shape, Cube):
            half = shape._side_length / 2
            return {"type": 2, "center": list(m.transform_point((0, 0, 0))), "radius": 0.0,
                    "normal": [0.0, 1.0, 0.0], "offset": 0.0, "extra": [half * s for s in axis_len], **md}
        raise NotImplementedError(f"taichi_export not supported for {type(shape).__name__}")


class Camera:
    """Pinhole camera; basis as scene/camera.py:32-34."""

    def __init__(self, position=(0, 0, 0), forward=(0, 0, -1), up=(0, 1, 0), fov=90, width=800, height=600):
        self.position = np.asarray(position, dtype=np.float64)
        self.forward = _unit(forward)
        self.right = _unit(np.cross(self.forward, _unit(up)))
        self.up = _unit(np.cross(self.right, self.forward))
        self.fov = float(fov)
        self.aspect_ratio = width / height if height != 0 else 1.0


class World:
    def __init__(self, background_color=(0.0, 0.0, 0.0), use_sky=False):
        self._objects = []
        self._cameras = []
        self._active_camera = None
        self._background_color = tuple(float(c) for c in background_color)
        self._use_sky = bool(use_sky)

    @property
    def objects(self):
        return list(self._objects)

    def add_object(self, obj):
        self._objects.append(obj)
        return obj

    def add_camera(self, cam):
        self._cameras.append(cam)
        if self._active_camera is None:
            self._active_camera = cam
        return cam

    @property
    def active_camera(self):
        return self._active_camera


class Image:
    """(H, W, 3) float32 display buffer (rendering/image.py:12-22)."""

    def __init__(self, width, height):
        self._width, self._height = int(width), int(height)
        self._pixels = np.zeros((self._height, self._width, 3), dtype=np.float32)

    width = property(lambda self: self._width)
    height = property(lambda self: self._height)
    pixels = property(lambda self: self._pixels)
    aspect_ratio = property(lambda self: self._width / self._height)
