"""ctypes binding of ``librxhip.so`` (the C ABI declared in include/roxy_hip.h).

The library is the product path: nothing here falls back to CPU code.  If
the library is missing, ``lib()`` raises with the command that builds it.
Status codes map to the exceptions the reference raises for the same
conditions: invalid arguments -> ValueError, capacity -> AssertionError
(rendering/taichi/bvh_builder.py:575-580 asserts), HIP failures ->
RuntimeError.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librxhip.so"

RX_OK = 0
RX_ERR_INVALID = 1
RX_ERR_CAPACITY = 2
RX_ERR_HIP = 3
RX_ERR_WORKSPACE = 4

RX_KIND_MISS, RX_KIND_PRIM, RX_KIND_BVH = 0, 1, 2
RX_PRIM_SPHERE, RX_PRIM_PLANE, RX_PRIM_CUBE, RX_PRIM_TRIANGLE = 0, 1, 2, 4
RX_MAT_DIFFUSE, RX_MAT_METAL, RX_MAT_DIELECTRIC, RX_MAT_EMISSIVE, RX_MAT_GLOSSY = 0, 1, 2, 3, 4
RX_MAX_LEAF_TRIS = 255
RX_MAX_TRIANGLES = 1 << 23
RX_MAX_BVH_DEPTH = 63
RX_FLAG_PROFILE = 1
RX_FLAG_COUNT_TRAFFIC = 2
RX_FLAG_PRIM_PREPASS = 4
RX_FLAG_SMALL_BLOCKS = 8
RX_FLAG_ADAPTIVE = 16
RX_FLAG_KEEP_QUEUE_ORDER = 32
RX_NOALLOC_SCRATCH_BYTES = 768 << 20
RX_SCENE_FINITE_BOXES = 1

NODE_BYTES = 64
TRI_BYTES = 48
SHADE_BYTES = 48
PRIM_BYTES = 48
MATERIAL_BYTES = 32
LIGHT_BYTES = 32
HEADER_BYTES = 32

vp = C.c_void_p
i32 = C.c_int32


class RxScene(C.Structure):
    _fields_ = [
        ("nodes", vp), ("header", vp), ("tris", vp), ("tri_shade", vp), ("materials", vp),
        ("prims", vp), ("prim_materials", vp), ("lights", vp), ("tri_uv", vp), ("textures", vp), ("texels", vp),
        ("n_nodes", i32), ("n_tris", i32), ("n_materials", i32), ("n_prims", i32), ("n_lights", i32),
        ("bvh_depth", i32), ("n_textures", i32), ("flags", i32),
    ]


class RxTexture(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("flip_v", i32), ("pad0", i32), ("offset", C.c_longlong),
                ("pad1", C.c_longlong)]


class RxCamera(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
        ("fov_tan", C.c_float), ("aspect", C.c_float), ("width", i32), ("height", i32),
    ]


class RxRenderParams(C.Structure):
    _fields_ = [
        ("max_depth", i32), ("use_sky", i32), ("direct_light_mode", i32), ("direct_light_max_depth", i32),
        ("sample_clamp", C.c_float), ("bg", C.c_float * 3), ("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32),
        ("flags", i32), ("samples_per_pass", i32), ("trav_blocks_per_cu", i32),
        ("adaptive_threshold", C.c_float), ("adaptive_min_samples", i32), ("adaptive_check_interval", i32),
    ]


class RxAccum(C.Structure):
    _fields_ = [("color", vp), ("normal", vp), ("albedo", vp), ("depth", vp), ("display", vp)]


class RxCounters(C.Structure):
    _fields_ = [(n, C.c_ulonglong) for n in (
        "rays", "shadow_rays", "node_pops", "tri_tests", "prim_tests",
        "shadow_node_pops", "shadow_tri_tests", "shadow_prim_tests")]


class RxProfile(C.Structure):
    _fields_ = [
        ("ms_generate", C.c_float), ("ms_extend", C.c_float), ("ms_shade", C.c_float),
        ("ms_connect", C.c_float), ("ms_accumulate", C.c_float),
        ("n_extend", i32), ("n_connect", i32), ("n_passes", i32),
        ("ms_prim_scan", C.c_float), ("n_prim_scan", i32),
        ("ms_reorder", C.c_float), ("n_reorder", i32),
    ]


class RxJob(C.Structure):
    _fields_ = [("scene", RxScene), ("camera", RxCamera), ("params", RxRenderParams), ("accum", RxAccum),
                ("first_sample", i32), ("samples_left", i32), ("path_offset", i32), ("pixel_offset", i32),
                ("rays", C.c_ulonglong), ("adaptive_error", C.c_float), ("adaptive_stopped", i32)]


class RxBvhRef(C.Structure):
    _fields_ = [("aabb_min", vp), ("aabb_max", vp), ("left", vp), ("right", vp), ("tri_start", vp),
                ("tri_count", vp)]


# name -> (restype, argtypes); every entry point of include/roxy_hip.h
SIGNATURES = {
    "rx_version": (C.c_int, []),
    "rx_last_error": (C.c_char_p, []),
    "rx_bvh_morton_shape": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "rx_bvh_workspace_size": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_size_t)]),
    "rx_bvh_build_morton": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, C.c_size_t, C.POINTER(RxBvhRef), vp, vp,
                                      vp, vp]),
    "rx_bvh_pack": (C.c_int, [C.POINTER(RxBvhRef), C.c_int, C.c_int, vp, vp, vp]),
    "rx_bvh_object_median_shape": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "rx_bvh_object_median_workspace_size": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_size_t)]),
    "rx_bvh_build_object_median": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, C.c_size_t, C.POINTER(RxBvhRef),
                                             vp, vp, vp, vp]),
    "rx_bvh_pack_triangles": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "rx_closest_hit": (C.c_int, [C.POINTER(RxScene), vp, C.c_int, vp, vp, vp, vp, vp, vp, vp]),
    "rx_any_hit": (C.c_int, [C.POINTER(RxScene), vp, vp, C.c_int, vp, vp, vp]),
    "rx_render_workspace_size": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_size_t)]),
    "rx_render_workspace_size_for_scene": (C.c_int, [C.POINTER(RxScene), C.c_int, C.c_int, C.c_int,
                                                     C.POINTER(C.c_size_t)]),
    "rx_render": (C.c_int, [C.POINTER(RxScene), C.POINTER(RxCamera), C.POINTER(RxRenderParams), C.c_int, C.c_int,
                            vp, C.c_size_t, C.POINTER(RxAccum), vp, C.POINTER(RxProfile), vp]),
    "rx_render_noalloc": (C.c_int, [C.POINTER(RxScene), C.POINTER(RxCamera), C.POINTER(RxRenderParams), C.c_int,
                                    C.c_int, C.POINTER(RxAccum), vp, vp]),
    "rx_batch_workspace_size": (C.c_int, [C.POINTER(RxJob), C.c_int, C.POINTER(C.c_size_t)]),
    "rx_batch_create": (C.c_int, [C.POINTER(RxJob), C.c_int, vp, C.c_size_t, C.c_int, C.POINTER(vp)]),
    "rx_batch_run": (C.c_int, [vp, C.c_int, vp, C.POINTER(RxProfile), vp]),
    "rx_batch_read_jobs": (C.c_int, [vp, C.POINTER(RxJob), vp]),
    "rx_radix_sort_scratch_size": (C.c_int, [C.c_longlong, C.POINTER(C.c_size_t)]),
    "rx_radix_sort_pairs": (C.c_int, [vp, vp, vp, vp, vp, C.c_longlong, C.c_int, C.c_int, vp, C.c_size_t, vp]),
    "rx_batch_graph_nodes": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "rx_batch_destroy": (C.c_int, [vp]),
    "rx_denoise": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                             C.c_float, C.c_float, vp, vp]),
    "rx_display": (C.c_int, [vp, C.c_int, vp, vp]),
    "rx_bvh_pack_uvs": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, vp]),
    "rx_flatten_indexed": (C.c_int, [vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, vp, vp,
                                     vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int]),
}

_LIB = None


class NativeLibraryMissing(RuntimeError):
    pass


def lib():
    """Load ``librxhip.so`` once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python -m roxy_dcc_amd.build` first")
        handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def last_error() -> str:
    msg = lib().rx_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = "") -> None:
    if status == RX_OK:
        return
    msg = last_error()
    text = f"{what}: {msg}" if what else msg
    if status == RX_ERR_INVALID:
        raise ValueError(text)
    if status == RX_ERR_CAPACITY:
        raise AssertionError(text)
    raise RuntimeError(f"librxhip status {status}: {text}")


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(stream=None) -> int | None:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def require_gpu():
    """The product path runs on the GPU only."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("roxy_dcc_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    lib()
