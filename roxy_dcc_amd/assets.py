"""Scene assets: OBJ/MTL meshes and Roxy Binary (RXB) mesh archives.

* ``MTLLoader`` / ``MTLMaterial`` — scene/io/mtl_loader.py:7-93: the .mtl
  fields the renderer uses and the same material mapping (emissive Ke,
  dielectric for d < 0.99, metal for illum 3-5, glossy for a non-zero Ks,
  diffuse otherwise; ``map_Kd`` becomes an ``ImageTexture`` path resolved
  against the .mtl's folder, last token of the line).
* ``OBJReader`` — the indexed path of scene/io/obj_reader.py:187-344:
  shared position/normal/uv arrays, fan triangulation of n-gons, 1-based
  and negative indices, ``o``/``g`` records as hierarchy nodes and one
  ``IndexedMesh`` leaf per run of faces under one ``usemtl``; group ids
  index one list of material names for the whole file.
* ``save_rxb_meshes`` / ``load_rxb_mesh`` / ``list_rxb_meshes`` /
  ``load_rxb_mesh_ref`` — scene/io/roxy_binary.py:13-255: an uncompressed
  numpy archive of index tables plus shared vertex buffers stored once,
  and a JSON metadata record (format "rxb", version 1 or 2).  Loaded with
  ``allow_pickle=False``.
"""
from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .world import Dielectric, Diffuse, Emissive, Glossy, ImageTexture, IndexedMesh, Metal, SceneObject, Shape

# ---------------------------------------------------------------- MTL


@dataclass
class MTLMaterial:
    name: str
    ka: tuple = None
    kd: tuple = None
    ks: tuple = None
    ke: tuple = None
    tf: tuple = None
    ns: float = 10.0
    ni: float = 1.0
    d: float = 1.0
    tr: float = 0.0
    illum: int = 1
    map_kd: str = None

    def to_material(self):
        """scene/io/mtl_loader.py:22-37."""
        texture = ImageTexture(self.map_kd) if self.map_kd else None
        if self.ke and any(c > 0 for c in self.ke):
            return Emissive(self.ke, intensity=1.0, albedo_texture=texture)
        if self.d < 0.99:
            return Dielectric(ior=self.ni, albedo=self.tf if self.tf else (1.0, 1.0, 1.0), albedo_texture=texture)
        if self.illum in (3, 4, 5):
            return Metal(self.kd or (1.0, 1.0, 1.0), roughness=max(0.0, 1.0 - self.ns / 1000.0),
                         albedo_texture=texture)
        if self.ks and any(c > 0 for c in self.ks):
            return Glossy(self.kd or (1.0, 1.0, 1.0), roughness=max(0.0, 1.0 - self.ns / 1000.0),
                          albedo_texture=texture)
        return Diffuse(self.kd or (0.8, 0.8, 0.8), albedo_texture=texture)


class MTLLoader:
    @staticmethod
    def load(path) -> dict:
        """Parse a .mtl file into {name: MTLMaterial} (mtl_loader.py:40-93)."""
        base_dir = Path(path).parent
        materials = []
        cur = None
        colours = {"Ka": "ka", "Kd": "kd", "Ks": "ks", "Ke": "ke", "Tf": "tf"}
        with open(path, "r") as f:
            for line in f:
                if line.startswith("#"):
                    continue
                parts = line.rsplit("#")[0].split()
                if not parts:
                    continue
                key = parts[0]
                if key == "newmtl":
                    if cur:
                        materials.append(cur)
                    cur = MTLMaterial(name=parts[1])
                elif cur is None:
                    continue
                elif key == "Ns":
                    cur.ns = float(parts[1])
                elif key == "Ni":
                    cur.ni = float(parts[1])
                elif key == "d":
                    cur.d = float(parts[1])
                elif key == "Tr":
                    cur.tr = float(parts[1])
                    cur.d = 1.0 - cur.tr
                elif key == "illum":
                    cur.illum = int(parts[1])
                elif key in ("map_kd", "map_Kd"):
                    cur.map_kd = str(base_dir / parts[-1])
                elif key in colours:
                    r, g, b = parts[1:4]
                    setattr(cur, colours[key], (float(r), float(g), float(b)))
        if cur:
            materials.append(cur)
        return {m.name: m for m in materials}


# ---------------------------------------------------------------- OBJ


class OBJReader:
    """OBJ reader producing IndexedMesh leaves (obj_reader.py:187-344)."""

    @staticmethod
    def _parse(path):
        file_path = Path(path)
        positions, normals, uvs = [], [], []
        materials = {}
        group_names, group_to_idx = [], {}
        events = []            # ("node", kind, name) or ("face", pos, uv, nrm, mtl)
        current_mtl = "default"

        def gid(name):
            if name not in group_to_idx:
                group_to_idx[name] = len(group_names)
                group_names.append(name)
            return group_to_idx[name]

        gid(current_mtl)

        def idx(raw, n):
            i = int(raw)
            return i - 1 if i > 0 else n + i

        with open(file_path, "r") as f:
            for line in f:
                parts = line.split()
                if not parts or parts[0].startswith("#"):
                    continue
                tok = parts[0]
                if tok == "mtllib":
                    mtl_path = file_path.parent / parts[1]
                    if mtl_path.exists():
                        materials.update({n: m.to_material() for n, m in MTLLoader.load(str(mtl_path)).items()})
                elif tok == "v":
                    positions.append([float(parts[1]), float(parts[2]), float(parts[3])])
                elif tok == "vn":
                    normals.append([float(parts[1]), float(parts[2]), float(parts[3])])
                elif tok == "vt":
                    uvs.append([float(parts[1]), float(parts[2])])
                elif tok in ("o", "g"):
                    events.append(("node", tok, parts[1] if len(parts) > 1 else ("object" if tok == "o" else "group")))
                elif tok == "usemtl":
                    new = parts[1] if len(parts) > 1 else "default"
                    if new != current_mtl:
                        events.append(("mtl", new))
                    current_mtl = new
                    gid(current_mtl)
                elif tok == "f":
                    vi, ti, ni = [], [], []
                    for part in parts[1:]:
                        vals = part.split("/")
                        vi.append(idx(vals[0], len(positions)))
                        ti.append(idx(vals[1], len(uvs)) if len(vals) > 1 and vals[1] else -1)
                        ni.append(idx(vals[2], len(normals)) if len(vals) > 2 and vals[2] else -1)
                    for i in range(1, len(vi) - 1):
                        events.append(("face", (vi[0], vi[i], vi[i + 1]), (ti[0], ti[i], ti[i + 1]),
                                       (ni[0], ni[i], ni[i + 1]), current_mtl))
        pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        nrm = np.asarray(normals, dtype=np.float64).reshape(-1, 3) if normals else None
        uv = np.asarray(uvs, dtype=np.float64).reshape(-1, 2) if uvs else None
        return pos, nrm, uv, materials, group_names, group_to_idx, events

    @staticmethod
    def load_indexed(path, name=None) -> SceneObject:
        """SceneObject hierarchy: ``o`` nodes under the root, ``g`` nodes under
        the current object, one IndexedMesh leaf per run of faces with one
        material under the current group node."""
        pos, nrm, uv, materials, group_names, group_to_idx, events = OBJReader._parse(path)
        root = SceneObject(name=name or Path(path).stem)
        obj_node = group_node = root
        run = {"pos": [], "uv": [], "nrm": [], "gid": [], "groups": set()}
        mtl = "default"

        def flush():
            if not run["pos"]:
                return
            mesh = IndexedMesh(pos, run["pos"], normals=nrm, tri_normal_idx=run["nrm"], uvs=uv,
                               tri_uv_idx=run["uv"], groups=group_names, tri_group_idx=run["gid"], name=mtl)
            mats = {g: materials.get(g, Diffuse((0.8, 0.8, 0.8))) for g in (run["groups"] or {"default"})}
            group_node.add_child(SceneObject(name=mtl, shapes=[Shape(mesh, None, mats)]))
            for k in ("pos", "uv", "nrm", "gid"):
                run[k] = []
            run["groups"] = set()

        for ev in events:
            if ev[0] == "node":
                flush()
                node = SceneObject(name=ev[2])
                if ev[1] == "o":
                    root.add_child(node)
                    obj_node = group_node = node
                else:
                    obj_node.add_child(node)
                    group_node = node
            elif ev[0] == "mtl":
                flush()
                mtl = ev[1]
            else:
                _, p3, t3, n3, m = ev
                mtl = m
                run["pos"].append(p3)
                run["uv"].append(t3)
                run["nrm"].append(n3)
                run["gid"].append(group_to_idx[m])
                run["groups"].add(m)
        flush()
        return root

    @staticmethod
    def load_as_indexed_mesh(path) -> IndexedMesh:
        """One flat IndexedMesh of every face; groups are the material names
        in first-use order after "default"."""
        pos, nrm, uv, _, group_names, group_to_idx, events = OBJReader._parse(path)
        faces = [ev for ev in events if ev[0] == "face"]
        return IndexedMesh(pos, [f[1] for f in faces], normals=nrm, tri_normal_idx=[f[3] for f in faces], uvs=uv,
                           tri_uv_idx=[f[2] for f in faces], groups=group_names,
                           tri_group_idx=[group_to_idx[f[4]] for f in faces], name=Path(path).stem)


# ---------------------------------------------------------------- RXB

RXB_FORMAT_VERSION = 2
_METADATA_KEY = "__metadata__"


def save_rxb_meshes(path, meshes):
    """Named IndexedMesh payloads into an RXB archive (roxy_binary.py:13-81);
    a vertex buffer shared by several meshes is stored once."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    items = list(meshes.items()) if hasattr(meshes, "items") else list(meshes)
    meta = {"format": "rxb", "version": RXB_FORMAT_VERSION, "buffers": {}, "meshes": {}}
    arrays, lookup = {}, {}

    def shared(kind, value):
        if value is None:
            return None
        a = np.asarray(value, dtype=np.float64)
        key = (kind, a.__array_interface__["data"][0], a.shape, str(a.dtype), a.strides)
        if key in lookup:
            return lookup[key]
        bid = f"{kind}_{sum(1 for b in meta['buffers'].values() if b['kind'] == kind)}"
        arrays[f"buffer.{bid}"] = a
        meta["buffers"][bid] = {"kind": kind, "array": f"buffer.{bid}", "dtype": str(a.dtype), "shape": list(a.shape)}
        lookup[key] = bid
        return bid

    for i, (name, mesh) in enumerate(items):
        if not isinstance(mesh, IndexedMesh):
            raise TypeError(f"RXB mesh payload {name!r} is not an IndexedMesh")
        prefix = f"mesh_{i}"
        meta["meshes"][name] = {
            "prefix": prefix, "name": mesh._name, "groups": list(mesh._groups),
            "has_normals": mesh._normals is not None, "has_uvs": mesh._uvs is not None,
            "positions_buffer": shared("positions", mesh._positions),
            "normals_buffer": shared("normals", mesh._normals), "uvs_buffer": shared("uvs", mesh._uvs),
            "vertex_count": mesh.vertex_count, "triangle_count": mesh.triangle_count,
        }
        for k in ("tri_pos_idx", "tri_normal_idx", "tri_uv_idx", "tri_group_idx"):
            arrays[f"{prefix}.{k}"] = np.asarray(getattr(mesh, "_" + k), dtype=np.int32)
    arrays[_METADATA_KEY] = np.frombuffer(json.dumps(meta, separators=(",", ":"), sort_keys=True).encode(),
                                          dtype=np.uint8)
    with open(path, "wb") as f:
        np.savez(f, **arrays)


def _metadata(archive, path):
    meta = json.loads(bytes(np.asarray(archive[_METADATA_KEY], dtype=np.uint8)).decode("utf-8"))
    if meta.get("format") != "rxb":
        raise ValueError(f"{path} is not a Roxy Binary file")
    if meta.get("version") not in (1, RXB_FORMAT_VERSION):
        raise ValueError(f"Unsupported RXB version {meta.get('version')} in {path}; expected {RXB_FORMAT_VERSION}")
    return meta


def load_rxb_metadata(path):
    with np.load(path, allow_pickle=False) as archive:
        return _metadata(archive, path)


def list_rxb_meshes(path):
    return list(load_rxb_metadata(path).get("meshes", {}).keys())


def load_rxb_mesh(path, name, cache=None):
    """One IndexedMesh from an RXB archive (roxy_binary.py:94-141); shared
    buffers are read once per (file, buffer) when a cache dict is given."""
    path = Path(path)
    with np.load(path, allow_pickle=False) as archive:
        meta = _metadata(archive, path)
        try:
            mm = meta["meshes"][name]
        except KeyError as exc:
            raise KeyError(f"RXB mesh {name!r} not found in {path}; available: "
                           f"{', '.join(meta['meshes'].keys())}") from exc
        prefix = mm["prefix"]

        def buffer(bid):
            bm = meta.get("buffers", {}).get(bid)
            if bm is None:
                raise KeyError(f"RXB buffer {bid!r} not found in {path}")
            key = (str(path.resolve()), bid)
            if cache is not None and key in cache:
                return cache[key]
            a = np.array(archive[bm["array"]], copy=True)
            if cache is not None:
                cache[key] = a
            return a

        def optional(legacy, key):
            if not mm.get(f"has_{legacy}"):
                return None
            return buffer(mm[key]) if mm.get(key) else np.array(archive[f"{prefix}.{legacy}"], copy=True)

        positions = buffer(mm["positions_buffer"]) if mm.get("positions_buffer") else np.array(
            archive[f"{prefix}.positions"], copy=True)
        return IndexedMesh(positions, np.array(archive[f"{prefix}.tri_pos_idx"], copy=True),
                           normals=optional("normals", "normals_buffer"),
                           tri_normal_idx=np.array(archive[f"{prefix}.tri_normal_idx"], copy=True),
                           uvs=optional("uvs", "uvs_buffer"),
                           tri_uv_idx=np.array(archive[f"{prefix}.tri_uv_idx"], copy=True),
                           groups=mm.get("groups") or ["default"],
                           tri_group_idx=np.array(archive[f"{prefix}.tri_group_idx"], copy=True),
                           name=mm.get("name", name))


def split_rxb_mesh_ref(ref, base_dir=None):
    """'file.rxb:meshes/name' -> (path, name) (roxy_binary.py:149-161)."""
    if ":" not in ref:
        raise ValueError(f"RXB mesh references must look like file.rxb:meshes/name: {ref}")
    path_text, mesh_path = ref.rsplit(":", 1)
    name = mesh_path[len("meshes/"):] if mesh_path.startswith("meshes/") else mesh_path
    if not name:
        raise ValueError(f"RXB mesh reference is missing a mesh name: {ref}")
    path = Path(path_text)
    if not path.is_absolute() and base_dir is not None:
        path = Path(base_dir) / path
    return path, name


def load_rxb_mesh_ref(ref, base_dir=None, cache=None):
    path, name = split_rxb_mesh_ref(ref, base_dir=base_dir)
    return load_rxb_mesh(path, name, cache=cache)
