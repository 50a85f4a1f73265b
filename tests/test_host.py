"""Host-side logic and the C ABI surface (no GPU needed)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import bvh_oracle
from oracle import denoise_oracle
from roxy_dcc_amd import _native as nat
from roxy_dcc_amd import scenes
from roxy_dcc_amd.bvh import GPUBVHBuilder, TriangleBatch
from roxy_dcc_amd.extract import flatten_scene
from roxy_dcc_amd.render_stats import RenderStats
from roxy_dcc_amd.renderer import TaichiRenderer, default_samples_per_pass

ROOT = Path(__file__).resolve().parent.parent


# ------------------------------------------------------------------ C ABI
def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "roxy_hip.h").read_text()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(rx_\w+)\s*\(", header, re.M))
    assert declared == set(nat.SIGNATURES)
    lib = nat.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.rx_version() == 6


def test_scene_workspace_holds_the_eye_relative_copies():
    """rx_render_workspace_size_for_scene = the render's own workspace plus
    the eye-relative copies: 64 bytes per node, 48 per triangle, 32 per
    primitive and the root record, and for 1..256 primitives 32 bytes of
    primitive masks per 8x8 tile of the camera (each array 256-byte
    aligned)."""
    lib = nat.lib()
    for n_nodes, n_tris, n_prims, depth in ((1000, 2000, 5, 10), (0, 0, 67, 0), (524287, 2_000_000, 67, 18),
                                            (1000, 2000, 0, 10), (1000, 2000, 300, 10)):
        sc = nat.RxScene()
        sc.n_nodes, sc.n_tris, sc.n_prims, sc.bvh_depth = n_nodes, n_tris, n_prims, depth
        a, b = C.c_size_t(), C.c_size_t()
        assert lib.rx_render_workspace_size_for_scene(C.byref(sc), 250, 128, 4, C.byref(a)) == 0
        assert lib.rx_render_workspace_size(250, 128, 4, depth, C.byref(b)) == 0
        extra = a.value - b.value
        tiles = 32 * 16 if 1 <= n_prims <= 256 else 0
        payload = 64 * n_nodes + 48 * n_tris + 32 * n_prims + 32 + 32 * tiles
        assert payload <= extra <= payload + 6 * 256, (extra, payload)
    assert lib.rx_render_workspace_size_for_scene(None, 256, 128, 4, C.byref(a)) == nat.RX_ERR_INVALID


def test_scene_workspace_holds_the_link_free_tree_copy():
    """With RX_SCENE_FINITE_BOXES and a tree whose heap slots (2^(depth+1) - 1)
    are at most four per node, the scene part holds the link-free copy of
    the tree instead of the eye-relative node records: 64 bytes per interior
    heap slot (2^depth - 1) in world space and again minus the eye, 4 bytes
    per heap slot for the leaves.  Without the flag, or for a tree too far
    from balanced, the scene part is as before."""
    lib = nat.lib()
    for n_nodes, n_tris, n_prims, depth, usable in ((524287, 2_000_000, 67, 18, True), (1000, 2000, 5, 10, True),
                                                    (1000, 2000, 5, 12, False), (3, 2, 0, 1, True)):
        for flags in (0, nat.RX_SCENE_FINITE_BOXES):
            sc = nat.RxScene()
            sc.n_nodes, sc.n_tris, sc.n_prims, sc.bvh_depth, sc.flags = n_nodes, n_tris, n_prims, depth, flags
            a, b = C.c_size_t(), C.c_size_t()
            assert lib.rx_render_workspace_size_for_scene(C.byref(sc), 250, 128, 4, C.byref(a)) == 0
            assert lib.rx_render_workspace_size(250, 128, 4, depth, C.byref(b)) == 0
            extra = a.value - b.value
            tiles = 32 * 16 if 1 <= n_prims <= 256 else 0
            base = 48 * n_tris + 32 * n_prims + 32 + 32 * tiles
            if flags and usable:
                interior = (1 << depth) - 1
                payload = base + 2 * 64 * interior + 4 * (2 * interior + 1)
            else:
                payload = base + 64 * n_nodes
            assert payload <= extra <= payload + 8 * 256, (n_nodes, depth, flags, extra, payload)


def test_morton_shape_closed_form_matches_recursion():
    for n in list(range(1, 40)) + [69_630, 100_003, 2_000_000]:
        for leaf in (1, 4, 8):
            a, b = C.c_int(), C.c_int()
            assert nat.lib().rx_bvh_morton_shape(n, leaf, C.byref(a), C.byref(b)) == 0
            assert (a.value, b.value) == bvh_oracle.closed_form_shape(n, leaf), (n, leaf)


def test_object_median_shape_matches_the_oracle_tree():
    """core/bvh.py's topology depends only on the count (the split is at
    len // 2 whatever the geometry): the library's shape equals the
    oracle's full build over real triangles for small counts and root leaf
    sizes, and a count recursion of core/bvh.py:5-13 for large counts,
    where the depth cap (depth > 20 makes a leaf) takes over past 4 * 2^21
    triangles."""
    from functools import lru_cache

    from roxy_dcc_amd.bvh import object_median_shape

    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 4, 5, 9, 17, 100, 1001):
        v0 = rng.normal(size=(n, 3))
        v1 = v0 + 0.1 * rng.normal(size=(n, 3))
        v2 = v0 + 0.1 * rng.normal(size=(n, 3))
        for leaf in (1, 4, 16):
            t = bvh_oracle.build_object_median(v0, v1, v2, max_leaf=leaf)
            assert object_median_shape(n, leaf) == (t.n_nodes, t.depth), (n, leaf)

    @lru_cache(maxsize=None)
    def rec(count, depth, root_leaf):
        if count <= (root_leaf if depth == 0 else 4) or depth > 20:
            return 1, depth
        a = rec(count // 2, depth + 1, root_leaf)
        b = rec(count - count // 2, depth + 1, root_leaf)
        return 1 + a[0] + b[0], max(a[1], b[1])

    for n in (69_630, 2_000_000, 8_388_608, 8_388_609, 20_000_000):
        for leaf in (4, 100):
            assert object_median_shape(n, leaf) == rec(n, 0, leaf), (n, leaf)
    assert object_median_shape(20_000_000, 4) == (2 ** 22 - 1, 21)   # leaves of 9-10 at the cap
    a, b = C.c_int(), C.c_int()
    assert nat.lib().rx_bvh_object_median_shape(10, 0, C.byref(a), C.byref(b)) == nat.RX_ERR_INVALID
    ws = C.c_size_t()
    assert nat.lib().rx_bvh_object_median_workspace_size(nat.RX_MAX_TRIANGLES + 1, 4, C.byref(ws)) \
        == nat.RX_ERR_CAPACITY


def test_abi_errors_are_reported_without_touching_the_gpu():
    lib = nat.lib()
    a, b = C.c_int(), C.c_int()
    assert lib.rx_bvh_morton_shape(10, 0, C.byref(a), C.byref(b)) == nat.RX_ERR_INVALID
    assert "max_leaf" in nat.last_error()
    ws = C.c_size_t()
    assert lib.rx_bvh_workspace_size(nat.RX_MAX_TRIANGLES + 1, 8, C.byref(ws)) == nat.RX_ERR_CAPACITY
    with pytest.raises(AssertionError):
        nat.check(nat.RX_ERR_CAPACITY)
    assert lib.rx_render(None, None, None, 0, 1, None, 0, None, None, None, None) == nat.RX_ERR_INVALID
    sc = nat.RxScene()
    assert lib.rx_closest_hit(C.byref(sc), None, 5, None, None, None, None, None, None, None) == nat.RX_ERR_INVALID
    assert lib.rx_denoise(None, None, None, None, 4, 4, 1, 1, 0.1, 0.1, 0.1, 0.1, 1.0, None, None) \
        == nat.RX_ERR_INVALID
    with pytest.raises(ValueError):
        nat.check(lib.rx_bvh_morton_shape(-1, 8, C.byref(a), C.byref(b)))


def test_render_workspace_size_grows_with_paths():
    ws1, ws2 = C.c_size_t(), C.c_size_t()
    assert nat.lib().rx_render_workspace_size(64, 64, 1, 0, C.byref(ws1)) == 0
    assert nat.lib().rx_render_workspace_size(64, 64, 4, 0, C.byref(ws2)) == 0
    # 49 four-byte words per path, plus the fixed traversal spill area
    assert ws2.value - ws1.value >= 3 * 64 * 64 * 49 * 4
    assert ws1.value >= 64 * 64 * 49 * 4
    # deep trees spill more stack entries (passes of 2^19 path slots or
    # more; smaller passes always run the deep-stack shape)
    wss, wsd = C.c_size_t(), C.c_size_t()
    assert nat.lib().rx_render_workspace_size(1024, 1024, 1, 0, C.byref(wss)) == 0
    assert nat.lib().rx_render_workspace_size(1024, 1024, 1, 40, C.byref(wsd)) == 0
    assert wsd.value > wss.value
    ws_small_deep = C.c_size_t()
    assert nat.lib().rx_render_workspace_size(64, 64, 1, 40, C.byref(ws_small_deep)) == 0
    assert ws_small_deep.value == ws1.value


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors have the sizes and field offsets gcc gives the header."""
    import subprocess

    structs = {"rx_scene": nat.RxScene, "rx_camera": nat.RxCamera, "rx_render_params": nat.RxRenderParams,
               "rx_accum": nat.RxAccum, "rx_counters": nat.RxCounters, "rx_profile": nat.RxProfile,
               "rx_bvh_ref": nat.RxBvhRef, "rx_job": nat.RxJob}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "roxy_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    for cname, size in (("rx_node", 64), ("rx_bvh_header", 32), ("rx_tri", 48), ("rx_tri_shade", 48),
                        ("rx_prim", 48), ("rx_material", 32), ("rx_light", 32)):
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert int(got[f"{cname}.{field}"]) == getattr(py, field).offset, (cname, field)
    assert int(got["rx_node"]) == nat.NODE_BYTES and int(got["rx_tri"]) == nat.TRI_BYTES
    assert int(got["rx_bvh_header"]) == nat.HEADER_BYTES and int(got["rx_prim"]) == nat.PRIM_BYTES
    assert int(got["rx_material"]) == nat.MATERIAL_BYTES and int(got["rx_light"]) == nat.LIGHT_BYTES
    assert int(got["rx_tri_shade"]) == nat.SHADE_BYTES


# ------------------------------------------------------------------ TriangleBatch
def _tri(x, mat_type=0, albedo=None):
    return {"v0": [x, 0, 0], "v1": [x + 1, 0, 0], "v2": [x, 1, 0], "n0": [0, 0, 1], "n1": [0, 0, 1],
            "n2": [0, 0, 1], "mat_type": mat_type, "albedo": albedo or [1, 1, 1], "roughness": 0.0, "ior": 1.0,
            "emission": 0.0}


def test_triangle_batch_from_dicts_and_defaults():
    b = TriangleBatch.from_dicts([_tri(0), _tri(2)])
    assert b.triangle_count == 2 and b.mat_idx is None
    assert b.uv0.shape == (2, 2) and b.has_uv.tolist() == [0, 0] and b.texture_idx.tolist() == [-1, -1]
    assert TriangleBatch.from_dicts([]).triangle_count == 0


def test_triangle_batch_validation():
    with pytest.raises(ValueError, match="same length"):
        TriangleBatch(np.zeros((2, 3)), np.zeros((2, 3)), np.zeros((3, 3)), np.zeros((2, 3)), np.zeros((2, 3)),
                      np.zeros((2, 3)), [0, 0], np.zeros((2, 3)), [0, 0], [1, 1], [0, 0])
    with pytest.raises(ValueError, match="match triangle count"):
        TriangleBatch(np.zeros((2, 3)), np.zeros((2, 3)), np.zeros((2, 3)), np.zeros((2, 3)), np.zeros((2, 3)),
                      np.zeros((2, 3)), [0], np.zeros((1, 3)), [0], [1], [0])


def test_indexed_batch_keeps_compact_palette():
    # tests/bvh_builder_tests.py:78-108
    arrays = {"v0": [[0, 0, 0], [2, 0, 0], [4, 0, 0]], "v1": [[1, 0, 0], [3, 0, 0], [5, 0, 0]],
              "v2": [[0, 1, 0], [2, 1, 0], [4, 1, 0]], "n0": [[0, 0, 1]] * 3, "n1": [[0, 0, 1]] * 3,
              "n2": [[0, 0, 1]] * 3, "uv0": [[0, 0]] * 3, "uv1": [[1, 0]] * 3, "uv2": [[0, 1]] * 3,
              "has_uv": [1, 1, 1], "group_idx": [0, 1, 0]}
    mats = {0: {"mat_type": 0, "albedo": [1, 0, 0], "roughness": 0, "ior": 1, "emission": 0, "texture_idx": 3},
            1: {"mat_type": 1, "albedo": [0, 1, 0], "roughness": 0.25, "ior": 1, "emission": 0, "texture_idx": -1},
            2: {"mat_type": 0, "albedo": [0, 0, 1], "roughness": 0, "ior": 1, "emission": 0, "texture_idx": -1}}
    batch = TriangleBatch.from_indexed_arrays(arrays, mats)
    assert batch.triangle_count == 3 and len(batch.mat_palette) == 2 and len(batch.mat_type) == 2
    assert list(batch.mat_idx) == [0, 1, 0] and list(batch.has_uv) == [1, 1, 1]
    assert batch.mat_palette[0]["texture_idx"] == 3


def test_concat_offsets_palettes():
    a = TriangleBatch.from_indexed_arrays(
        {"v0": [[0, 0, 0]], "v1": [[1, 0, 0]], "v2": [[0, 1, 0]], "n0": [[0, 0, 1]], "n1": [[0, 0, 1]],
         "n2": [[0, 0, 1]], "group_idx": [0]}, {0: {"mat_type": 0, "albedo": [1, 0, 0], "roughness": 0, "ior": 1,
                                                    "emission": 0}})
    c = TriangleBatch.concat([a, a])
    assert c.triangle_count == 2 and list(c.mat_idx) == [0, 1] and len(c.mat_palette) == 2


def test_builder_palette_dedup_first_appearance():
    # tests/bvh_builder_tests.py:20-33 (palette sizes), without a GPU build
    b = GPUBVHBuilder()
    assert len(set(b._build_material_indices(TriangleBatch.from_dicts([_tri(0), _tri(2)])))) == 1
    b = GPUBVHBuilder()
    idx = b._build_material_indices(TriangleBatch.from_dicts([_tri(0), _tri(2, albedo=[0.5, 0.5, 0.5]), _tri(3)]))
    assert idx.tolist() == [0, 1, 0] and len(b.mat_palette) == 2
    assert b.mat_palette[1]["albedo"] == [0.5, 0.5, 0.5]


def test_builder_longest_axis_ties():
    b = GPUBVHBuilder()
    assert b._longest_axis([0, 0, 0], [1, 1, 1]) == 0
    assert b._longest_axis([0, 0, 0], [1, 2, 2]) == 1
    assert b._longest_axis([0, 0, 0], [1, 1, 3]) == 2


# ------------------------------------------------------------------ renderer front end
def test_renderer_validation_and_defaults():
    world, *_ = scenes.cornell()
    with pytest.raises(ValueError, match="direct_light_mode"):
        TaichiRenderer(world, None, None, direct_light_mode="bogus")
    r = TaichiRenderer(world, None, None, max_depth=7, direct_light_max_depth=None)
    assert r._direct_light_max_depth == 7
    assert TaichiRenderer._direct_light_mode_to_id("all") == 1 and TaichiRenderer._direct_light_mode_to_id(0) == 0
    assert repr(r) == "TaichiRenderer(samples=16, max_depth=7, direct_light_mode='one')"


def test_default_samples_per_pass():
    assert default_samples_per_pass(256, 256, 16) == 16
    assert default_samples_per_pass(1920, 1080, 16) == 16
    assert default_samples_per_pass(1920, 1080, 64) == 16
    assert default_samples_per_pass(512, 512, 256) == 128
    assert default_samples_per_pass(1920, 1080, 16, paths=1 << 21) == 1
    assert default_samples_per_pass(512, 512, 64, paths=1 << 21) == 8


def test_render_stats_report():
    s = RenderStats(4, 2, 16, 16, 5, 1000, 2.0, primitive_count=3, bvh_nodes=15)
    assert s.rays_per_pixel == 125.0 and s.rays_per_second == 500.0
    rep = s.format_report()
    assert "rays cast:       1,000" in rep and "BVH nodes:       15" in rep


# ------------------------------------------------------------------ scenes
def test_scene_generators_have_the_configured_sizes():
    flat = flatten_scene(scenes.cornell()[0])
    assert flat.triangle_count == 36 and len(flat.prims) == 0
    w, W, H, st = scenes.mesh_scene()
    flat = flatten_scene(w)
    assert flat.triangle_count == 10 + 20 * 59 * 59 and len(flat.prims.lights()) == 1
    assert (W, H, st["samples"], st["max_depth"]) == (512, 512, 64, 8)
    pos, idx = scenes.uv_sphere(50, 41)
    assert len(idx) == 4000
    w, W, H, st = scenes.large_scene(n_meshes=3)
    flat = flatten_scene(w)
    assert flat.triangle_count == 12_000 and len(flat.prims) == 67 and len(flat.prims.lights()) == 3


def test_cornell_walls_follow_the_configuration():
    flat = flatten_scene(scenes.cornell()[0])
    tb = flat.triangles()
    alb = np.asarray([tb.mat_palette[i]["albedo"] for i in tb.mat_idx], np.float32)
    red = np.all(np.isclose(alb, [0.65, 0.05, 0.05]), axis=1)
    assert red.sum() == 2 and np.all(tb.v0[red][:, 0] == 0)  # left wall is x = 0
    em = np.asarray([tb.mat_palette[i]["type"] for i in tb.mat_idx]) == 3
    assert em.sum() == 2 and np.allclose(tb.v0[em][:, 1], 0.999)


# ------------------------------------------------------------------ denoise oracle
def test_denoise_oracle_reference_cases():
    # tests/denoise_tests.py:7-107
    assert denoise_oracle.display(np.array([[[-1.0, 0.25, 4.0]]], np.float32)).tolist() == [[[0.0, 0.5, 1.0]]]
    p = np.zeros((5, 5, 3), np.float32)
    p[:, :] = 0.2
    p[2, 2] = 1.0
    f = denoise_oracle.denoise(p, radius=1, sigma_color=0.2, amount=0.8)
    assert f.shape == p.shape and np.isfinite(f).all() and f[2, 2, 0] > f[0, 0, 0]
    q = np.random.default_rng(1).random((4, 4, 3), dtype=np.float32)
    assert denoise_oracle.denoise(q, radius=0) is q and denoise_oracle.denoise(q, amount=0.0) is q
    b = np.zeros((3, 5, 3), np.float32)
    b[:, 2:] = 1.0
    plain = denoise_oracle.denoise(b, 1, 10.0, 1.0)
    nrm = np.zeros((3, 5, 3), np.float32)
    nrm[:, :2] = [0, 1, 0]
    nrm[:, 2:] = [0, -1, 0]
    dep = np.ones((3, 5), np.float32)
    dep[:, 2:] = 10.0
    alb = np.zeros((3, 5, 3), np.float32)
    alb[:, :2] = 0.1
    alb[:, 2:] = 0.9
    for kw in ({"normal": nrm, "sigma_normal": 0.1}, {"albedo": alb, "sigma_albedo": 0.05},
               {"depth": dep, "sigma_depth": 0.1}):
        g = denoise_oracle.denoise(b, 1, 10.0, 1.0, **kw)
        assert g[1, 1, 0] < plain[1, 1, 0] * 0.25 and g[1, 2, 0] > plain[1, 2, 0] + 0.20
    with pytest.raises(ValueError, match="normal guide"):
        denoise_oracle.denoise(b, normal=np.zeros((3, 4, 3), np.float32))


def test_denoise_oracle_fixed_point_and_dilation():
    u = np.full((9, 11, 3), 0.37, np.float32)
    np.testing.assert_allclose(denoise_oracle.denoise(u, 2, 0.1, 1.0, dilation=3), u, rtol=1e-6)
    # a dilated pass equals the dense filter on the sub-lattice when the
    # image is periodic-free and the taps stay inside
    img = np.random.default_rng(3).random((40, 40, 3), dtype=np.float32)
    d2 = denoise_oracle.denoise(img, 1, 0.3, 1.0, dilation=2)
    sub = img[10:31:2, 10:31:2]
    dense = denoise_oracle.denoise(sub, 1, 0.3, 1.0)
    np.testing.assert_allclose(d2[12:29:2, 12:29:2], dense[1:-1, 1:-1], rtol=1e-6)


# ------------------------------------------------------------------ serving (host logic)
def test_scene_cache_lru_respects_pins_and_budget():
    from roxy_dcc_amd.serve import SceneCache

    class Fake:
        def __init__(self, n):
            self.n = n

        def hbm_bytes(self):
            return self.n

    loads = []

    def loader(key):
        loads.append(key)
        return Fake(100), None

    c = SceneCache(loader, max_scenes=2, max_bytes=250)
    c.get("a")
    c.get("b")
    c.get("a")           # hit, refreshes a
    c.get("c")           # evicts b (least recent)
    assert list(c.entries) == ["a", "c"] and c.evictions == 1
    c.get("d", pinned=["a", "c"])   # over both budgets, but a and c are pinned
    assert set(c.entries) == {"a", "c", "d"}
    assert loads == ["a", "b", "c", "d"]


def test_poisson_arrivals_rate():
    from roxy_dcc_amd.serve import poisson_arrivals

    t = poisson_arrivals(50.0, 20000, seed=1)
    assert np.all(np.diff(t) > 0)
    assert abs(len(t) / t[-1] - 50.0) < 2.0


def test_autotune_candidates_and_apply():
    """Pass-count candidates stay within the path and memory budgets; a
    tuning sets every job's samples per pass and launch shape."""
    from types import SimpleNamespace

    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd.autotune import BYTES_PER_PATH, Tuning, pass_candidates

    jobs = [SimpleNamespace(width=w, height=w, samples=s, samples_per_pass=1, trav_blocks_per_cu=0, flags=0)
            for w, s in [(64, 4), (512, 32), (256, 17), (128, 9)]]
    ks = pass_candidates(jobs, budget_bytes=1 << 30)
    assert ks and all(1 <= k <= 32 for k in ks) and len(set(ks)) == len(ks)
    for k in ks:
        per_pass = sum(j.width * j.height * -(-j.samples // k) for j in jobs)
        assert per_pass * BYTES_PER_PATH <= 1 << 30
    t = Tuning(passes=4, blocks_per_cu=6, small_blocks=True, prim_prepass=True)
    t.apply(jobs)
    assert [j.samples_per_pass for j in jobs] == [1, 8, 5, 3]
    assert all(j.trav_blocks_per_cu == 6 for j in jobs)
    assert all(j.flags == nat.RX_FLAG_SMALL_BLOCKS | nat.RX_FLAG_PRIM_PREPASS for j in jobs)


def test_share_pass_water_filling():
    """Pass budget split: every job gets >= 1 sample and never more than it
    needs; small jobs are served in full; the budget is not exceeded once
    every job's one-sample minimum fits."""
    from roxy_dcc_amd.serve import share_pass

    demands = [(64 * 64, 4), (512 * 512, 32), (1920 * 1080, 64), (128 * 128, 16)]
    spp = share_pass(demands, 1 << 23, "fair")
    assert all(1 <= s <= need for s, (_, need) in zip(spp, demands))
    assert spp[0] == 4 and spp[3] == 16
    assert sum(s * px for s, (px, _) in zip(spp, demands)) <= 1 << 23
    # one job alone takes as much of the budget as it can use
    assert share_pass([(512 * 512, 64)], 1 << 23, "fair") == [32]
    # a frame larger than the budget still advances one sample per pass
    assert share_pass([(1920 * 1080, 8)], 1 << 20, "fair") == [1]
    # first come first served with backfill: the oldest job takes what it
    # can use, the rest advance at least one sample
    fifo = share_pass(demands, 1 << 23, "fifo")
    assert all(1 <= s <= need for s, (_, need) in zip(fifo, demands))
    assert fifo[0] == 4 and fifo[1] == 23 and fifo[2] == 1 and fifo[3] == 16
    assert sum(s * px for s, (px, _) in zip(fifo, demands)) <= 1 << 23
    # arrival order against remaining work: the large frame first in line
    late = [(1920 * 1080, 64), (512 * 512, 32)]
    assert share_pass(late, 1 << 23, "fifo") == [3, 8]
    assert share_pass(late, 1 << 23, "srpt") == [1, 24]
    assert share_pass(demands, 1 << 23, "srpt") == [4, 23, 1, 16]
    with pytest.raises(ValueError):
        share_pass(demands, 1 << 23, "lifo")


def test_share_pass_deadline_guard():
    """The default policy: shortest remaining work first unless a job would
    then finish later than first come first served (plus the slack) lets it;
    jobs whose next sample does not fit sit the pass out, the first in line
    always advances, and the budget is never exceeded otherwise."""
    from roxy_dcc_amd.serve import deadline_order, fifo_deadlines, share_pass

    demands = [(64 * 64, 4), (512 * 512, 32), (1920 * 1080, 64), (128 * 128, 16)]
    spp = share_pass(demands, 1 << 23)
    # the 512^2 job's last two samples overshoot the budget by 3%: it
    # finishes in this pass
    assert spp == [4, 32, 0, 16]
    assert sum(s * px for s, (px, _) in zip(spp, demands)) <= 1.25 * (1 << 23)
    assert share_pass([(512 * 512, 41)], 1 << 23) == [32]      # over by more than 25%: not stretched
    assert share_pass([(512 * 512, 36)], 1 << 23) == [36]
    # a long job that arrived first: shortest-first while the slack allows,
    # arrival order once the short job would push it past its deadline
    work = [1920 * 1080 * 64, 512 * 512 * 32]
    assert deadline_order(work, fifo_deadlines(work, work[1])) == [1, 0]
    assert deadline_order(work, fifo_deadlines(work, work[1] - 1)) == [0, 1]
    assert deadline_order(work, fifo_deadlines(work, 0)) == [0, 1]
    # a stream of small jobs behind a large one cannot overtake it for good
    work = [1000] + [10] * 200
    order = deadline_order(work, fifo_deadlines(work, 100))
    assert order.index(0) <= 10
    # deadlines that no order meets: earliest deadline first
    assert deadline_order([5, 5, 5], [4, 3, 20]) == [1, 0, 2]
    # explicit deadlines and costs: the costly job is the long one
    spp = share_pass([(100, 10), (100, 10)], 10 ** 6, deadlines=[10 ** 9, 10 ** 9], costs=[5.0, 1.0])
    assert spp == [10, 10]
    # every order meets its deadlines
    rng = np.random.default_rng(3)
    for _ in range(200):
        w = rng.integers(1, 1000, size=int(rng.integers(1, 12))).tolist()
        slack = int(rng.integers(0, 500))
        order = deadline_order(w, fifo_deadlines(w, slack))
        assert sorted(order) == list(range(len(w)))
        t, dl = 0, np.cumsum(w) + slack
        for k in order:
            t += w[k]
            assert t <= dl[k]
    # a frame larger than the budget still advances one sample per pass
    assert share_pass([(1920 * 1080, 8)], 1 << 20) == [1]


def test_server_loop_serves_every_request_once(monkeypatch):
    """RenderServer.run's bookkeeping with a stand-in for the device batch
    (host arrays, one pass = every job advances by its share) and a clock
    that advances 1 ms per reading: every request finishes exactly once,
    with all its samples and its own frame, across idle gaps between
    arrivals; frames of the last pass are read back after the loop drains."""
    import types
    import contextlib
    import torch
    from roxy_dcc_amd import serve

    class Job:
        def __init__(self, req):
            self.width, self.height, self.samples = req.width, req.height, req.samples
            self.first_sample, self.samples_per_pass, self.tag = 0, 1, req
            self.frame = types.SimpleNamespace(color=torch.full((req.height, req.width, 3), float(req.rid)))

    class FakeBatch:
        created = 0

        def __init__(self, jobs, use_graph=False):
            FakeBatch.created += 1
            self.jobs, self.first, self.rays = jobs, [j.first_sample for j in jobs], [0] * len(jobs)

        def run(self, passes):
            for i, j in enumerate(self.jobs):
                take = min(j.samples - self.first[i], j.samples_per_pass)
                self.first[i] += take
                self.rays[i] += 2 * j.width * j.height * take

        def read_jobs(self):
            return [types.SimpleNamespace(rays=r, first_sample=f, samples_left=j.samples - f)
                    for j, f, r in zip(self.jobs, self.first, self.rays)]

        def close(self):
            pass

    monkeypatch.setattr(serve, "Batch", FakeBatch)
    monkeypatch.setattr(serve, "request_job", lambda req, scene, cam, opts, samples_per_pass: Job(req))
    monkeypatch.setattr(torch.cuda, "Stream", lambda: None)
    monkeypatch.setattr(torch.cuda, "stream", lambda s: contextlib.nullcontext())
    ticks = iter(range(10 ** 6))
    clock = lambda: next(ticks) * 1e-3
    sizes = [(64, 64, 8), (512, 512, 64), (32, 32, 4), (1920, 1080, 16), (128, 128, 32), (16, 16, 1)]
    arrivals = [0.0, 0.0, 0.003, 0.004, 0.5, 0.501]          # an idle gap before the last two
    reqs = [serve.Request(i, ("s", i % 2), w, h, n, a) for i, ((w, h, n), a) in enumerate(zip(sizes, arrivals))]
    scene = types.SimpleNamespace(hbm_bytes=lambda: 1 << 20)
    cache = serve.SceneCache(lambda key: (scene, object(), {}), max_scenes=2)
    server = serve.RenderServer(cache, max_jobs=4, pass_paths=1 << 20)
    done = server.run(reqs, clock=clock, idle_sleep=0.0)
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
    for r in done:
        assert r.samples_done == r.samples and r.rays == 2 * r.width * r.height * r.samples
        assert r.frame.shape == (r.height, r.width, 3) and (r.frame == r.rid).all()
        assert r.finished >= r.arrival and r.admitted >= r.arrival
    assert server.rays_done == sum(r.rays for r in done)
    assert server.passes >= 16 and FakeBatch.created >= 2     # the 1080p job alone takes 16 passes of 2^20


def test_share_pass_join_rule():
    """With ``join`` (a pass's fixed cost, in paths), a job joins a pass that
    already holds jobs only while (jobs in the pass) x (its paths) is at most
    ``join`` x (it and the jobs after it in the order): small jobs ride
    along, a large one waits for its own pass unless many jobs queue behind
    it; the order and the samples of the jobs that join are unchanged."""
    from roxy_dcc_amd.serve import share_pass

    far = [1e12] * 3
    three = [(960 * 540, 8), (960 * 540, 32), (512 * 512, 64)]
    assert share_pass(three, 1 << 25, deadlines=far) == [8, 32, 64]
    # 16.6M paths after a 4.1M-path job, two jobs from the end: 16.6M > 2 x 4.2M
    assert share_pass(three, 1 << 25, deadlines=far, join=1 << 22) == [8, 0, 0]
    assert share_pass(three, 1 << 25, deadlines=far, join=1 << 23) == [8, 32, 0]
    # small jobs join a 1080p job that its deadline puts first
    mixed = [(1920 * 1080, 16), (256 * 256, 4), (128 * 128, 16)]
    dl = [1920 * 1080 * 16, 1e12, 1e12]
    assert share_pass(mixed, 1 << 25, deadlines=dl) == [16, 4, 16]
    assert share_pass(mixed, 1 << 25, deadlines=dl, join=1 << 22) == [16, 4, 16]
    # a queue of ten equal 2M-path jobs: the k-th (from 0) joins while
    # k x 2M <= 4.2M x (10 - k), so seven share the pass
    ten = [(1024 * 1024, 2)] * 10
    assert share_pass(ten, 1 << 25, deadlines=[1e12] * 10) == [2] * 10
    assert share_pass(ten, 1 << 25, deadlines=[1e12] * 10, join=1 << 22) == [2] * 7 + [0] * 3
    # the first job always advances, whatever its size
    assert share_pass([(1920 * 1080, 64)], 1 << 25, join=1) == [16]


def _random_indexed_mesh(rng, n_pos=200, n_tris=500, normals=True, uvs=True):
    from roxy_dcc_amd.world import IndexedMesh

    pos = rng.normal(size=(n_pos, 3)) * 3.0
    tri = rng.integers(0, n_pos, size=(n_tris, 3)).astype(np.int32)
    tri[:5, 1] = tri[:5, 0]                       # degenerate: zero-area faces
    kw = {}
    if normals:
        nrm = rng.normal(size=(n_pos, 3))
        nrm[:3] = 0.0                             # zero normals stay unnormalised
        tni = rng.integers(0, n_pos, size=(n_tris, 3)).astype(np.int32)
        tni[::7, 2] = -1                          # these triangles fall back to the face normal
        kw.update(normals=nrm, tri_normal_idx=tni)
    if uvs:
        kw.update(uvs=rng.random((50, 2)), tri_uv_idx=np.where(rng.random((n_tris, 3)) < 0.9,
                                                               rng.integers(0, 50, (n_tris, 3)), -1).astype(np.int32))
    return IndexedMesh(pos, tri, **kw)


def _random_matrices(rng):
    from roxy_dcc_amd.world import Mat4

    m = Mat4.trs(translation=rng.normal(size=3) * 10, rotation_deg=rng.uniform(-180, 180, 3),
                 scale=rng.uniform(0.2, 3.0, 3))
    p = Mat4(m.rows.copy())
    p.rows[3] = [0.01, -0.02, 0.03, 1.0]          # a projective row exercises the divide
    return [(None, None), (m, m.inverse().transpose()), (p, p.inverse().transpose())]


def test_native_flatten_equals_restatement():
    """rx_flatten_indexed equals the fp64 numpy restatement of
    indexed_triangle_arrays bit for bit: affine and projective matrices,
    identity, per-vertex and face normals, missing uvs, degenerate faces."""
    from roxy_dcc_amd.extract import indexed_arrays_native

    rng = np.random.default_rng(17)
    for normals, uvs in [(True, True), (False, False), (True, False)]:
        mesh = _random_indexed_mesh(rng, normals=normals, uvs=uvs)
        for m, nm in _random_matrices(rng):
            a = mesh.indexed_triangle_arrays(m, nm)
            b = indexed_arrays_native(mesh, m, nm)
            for k in ("v0", "v1", "v2", "n0", "n1", "n2", "uv0", "uv1", "uv2", "has_uv", "group_idx"):
                assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), (k, normals, uvs)


def test_flatten_restatement_matches_reference_matmul_form():
    """The elementwise fp64 order agrees with the reference's matmul form
    (scene/mesh.py:547-568: flat @ m[:3,:3].T + m[:3,3]) to within one fp32
    ulp, and exactly for almost every coordinate."""
    from roxy_dcc_amd.world import _xf_points, _xf_vectors

    rng = np.random.default_rng(5)
    pts = rng.normal(size=(20000, 3)) * 50
    for m, nm in _random_matrices(rng)[1:]:
        r = m.rows
        ref = pts @ r[:3, :3].T + r[:3, 3]
        w = pts @ r[3, :3].T + r[3, 3]
        div = np.abs(w) > 1e-12
        ref[div] /= w[div, None]
        got = _xf_points(pts, m)
        a, b = ref.astype(np.float32), got.astype(np.float32)
        assert np.all(np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64)) <= 1)
        assert np.mean(a == b) > 0.999
        vref = (pts @ nm.rows[:3, :3].T).astype(np.float32)
        vgot = _xf_vectors(pts, nm).astype(np.float32)
        assert np.all(np.abs(vref.view(np.int32).astype(np.int64) - vgot.view(np.int32).astype(np.int64)) <= 1)


def test_native_flatten_rejects_bad_indices():
    import pytest

    from roxy_dcc_amd.extract import indexed_arrays_native
    from roxy_dcc_amd.world import IndexedMesh

    mesh = IndexedMesh(np.zeros((3, 3)), np.array([[0, 1, 3]], np.int32))
    with pytest.raises(ValueError):
        indexed_arrays_native(mesh)


def test_texture_cache_dedup_and_capacity():
    """extractor.py:93-117: one cache slot per distinct texture (path, flip_v,
    identity of an in-memory texture) in first-use order; more than
    MAX_TEXTURES raises ValueError; u8 texels as scene/textures.py:50-52."""
    import pytest

    from roxy_dcc_amd.extract import MAX_TEXTURES, flatten_scene
    from roxy_dcc_amd.world import Diffuse, ImageTexture, IndexedMesh, Mat4, SceneObject, Shape, World

    def world_with(textures):
        w = World()
        for i, t in enumerate(textures):
            mesh = IndexedMesh(np.eye(3), np.array([[0, 1, 2]], np.int32), uvs=np.zeros((3, 2)),
                               tri_uv_idx=np.array([[0, 1, 2]], np.int32))
            w.add_object(SceneObject(f"o{i}", [Shape(mesh, Diffuse((1, 1, 1), albedo_texture=t))], Mat4()))
        return w

    a = ImageTexture.from_array(np.zeros((2, 2, 3)))
    b = ImageTexture.from_array(np.ones((2, 2, 3)))
    flat = flatten_scene(world_with([a, b, a, None]))
    assert flat.textures == [a, b]
    assert list(flat.triangles().texture_idx) == [0, 1, 0, -1]
    with pytest.raises(ValueError):
        flatten_scene(world_with([ImageTexture.from_array(np.zeros((1, 1, 3))) for _ in range(MAX_TEXTURES + 1)]))
    px = ImageTexture.from_array(np.array([[[0.0, 0.5, 1.0]]])).load_pixels_u8()
    assert px.dtype == np.uint8 and px.tolist() == [[[0, 128, 255]]]
    assert ImageTexture.from_array(np.array([[[0, 128, 255]]])).load_pixels_u8().tolist() == [[[0, 128, 255]]]
