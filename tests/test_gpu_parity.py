"""GPU parity against the CPU oracle, through the C ABI.

Closest hits: primitive ids, kinds, t, u, v must be bit-identical to the
oracle's restatement of rendering/taichi/intersect.py (the oracle keeps the
reference's node layout; the GPU uses packed 64-byte nodes).  BVH builds:
node arrays and leaf order bit-identical to the numpy restatement of
rendering/taichi/bvh_builder.py.  Renders: same Philox draws and fp32
operation order on both sides, so every pixel, guide buffer and ray count
must be identical.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import bvh_oracle
from oracle import denoise_oracle
from oracle import oracle as orc
from tests import helpers

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _gpu_build(v0, v1, v2, leaf=8, method=None):
    from roxy_dcc_amd.bvh import GPUBVHBuilder, TriangleBatch

    n = len(v0)
    z = np.zeros((n, 3), np.float32)
    batch = TriangleBatch(v0, v1, v2, z + [0, 0, 1], z + [0, 0, 1], z + [0, 0, 1], np.zeros(n, np.int32), z + 0.5,
                          np.zeros(n, np.float32), np.ones(n, np.float32), np.zeros(n, np.float32))
    b = GPUBVHBuilder()
    b.MAX_LEAF_SIZE = leaf
    b.build([batch], method=method)
    _torch().cuda.synchronize()
    return b


def _assert_tree_equal(b, t):
    r = b.reference_arrays()
    print(f"nodes gpu {b.node_count} oracle {t.n_nodes} depth {b.depth} / {t.depth}", flush=True)
    assert b.node_count == t.n_nodes
    assert b.depth == t.depth
    np.testing.assert_array_equal(b.ordered_tris, t.order)
    for k in ("left", "right", "tri_start", "tri_count"):
        np.testing.assert_array_equal(r[k], getattr(t, k), err_msg=k)
    for k in ("aabb_min", "aabb_max"):
        np.testing.assert_array_equal(r[k].view(np.uint32), getattr(t, k).view(np.uint32), err_msg=k)


@pytest.mark.parametrize("n,leaf,seed", [(1, 8, 0), (2, 8, 0), (8, 1, 1), (9, 4, 2), (36, 8, 3), (1000, 8, 4),
                                         (4097, 3, 5), (65536, 8, 6), (100003, 1, 7)])
def test_morton_build_matches_oracle(n, leaf, seed):
    rng = np.random.default_rng(seed)
    base = rng.uniform(-50, 50, size=(n, 3)).astype(np.float32)
    v0 = base
    v1 = (base + rng.normal(size=(n, 3))).astype(np.float32)
    v2 = (base + rng.normal(size=(n, 3))).astype(np.float32)
    b = _gpu_build(v0, v1, v2, leaf)
    _assert_tree_equal(b, bvh_oracle.build_morton(v0, v1, v2, leaf))


def test_morton_build_degenerate_axes_and_duplicates():
    # flat in z (extent 0 -> 1) and many identical centroids (stable ties)
    rng = np.random.default_rng(11)
    n = 5000
    v0 = np.zeros((n, 3), np.float32)
    v0[:, :2] = rng.integers(0, 4, size=(n, 2))
    v1 = v0 + np.float32([1, 0, 0])
    v2 = v0 + np.float32([0, 1, 0])
    b = _gpu_build(v0, v1, v2, 8)
    _assert_tree_equal(b, bvh_oracle.build_morton(v0, v1, v2, 8))


@pytest.mark.parametrize("name", ["cornell", "mesh"])
def test_scene_bvh_matches_oracle(name):
    from roxy_dcc_amd.extract import extract_scene

    world, *_ = helpers.load(name)
    ex = extract_scene(world)
    tb = ex.flat.triangles()
    _assert_tree_equal(ex.scene.builder, bvh_oracle.build_morton(tb.v0, tb.v1, tb.v2, 8))


def test_median_build_matches_oracle():
    rng = np.random.default_rng(3)
    n = 3000
    v0 = rng.uniform(0, 10, size=(n, 3)).astype(np.float32)
    v1 = (v0 + rng.uniform(0, 1, size=(n, 3))).astype(np.float32)
    v2 = (v0 + rng.uniform(0, 1, size=(n, 3))).astype(np.float32)
    b = _gpu_build(v0, v1, v2, 8, method="median")
    _assert_tree_equal(b, bvh_oracle.build_median(v0, v1, v2, 8))


@pytest.mark.parametrize("method", ["morton", "median"])
def test_builds_match_reference_builder_golden(method):
    """Both GPUBVHBuilder methods against the reference builder's own
    output (tests/golden/bvh_builder.npz): node arrays and leaf order bit
    for bit, on every seeded case and on the Cornell and mesh scenes."""
    from pathlib import Path

    from roxy_dcc_amd.extract import extract_scene
    from tests.golden import make_bvh_builder_golden as gb
    from tests.test_reference_goldens import _median_comparable, assert_tree_matches

    g = np.load(Path(__file__).resolve().parent / "golden" / "bvh_builder.npz")
    _median_comparable(g, method)

    def arrays(b):
        r = b.reference_arrays()
        return dict(r, order=b.ordered_tris)

    for k, (n, seed) in enumerate(g["cases"]):
        v0, v1, v2 = gb.mesh(int(n), int(seed), str(g["kinds"][k]))
        print(f"golden case {k}: n={n} {g['kinds'][k]} {method}", flush=True)
        assert_tree_matches(arrays(_gpu_build(v0, v1, v2, 8, method=method)), g, f"case{k}_{method}")
    for name in ("cornell", "mesh"):
        world, *_ = helpers.load(name)
        ex = extract_scene(world, method=method)
        a = arrays(ex.scene.builder)
        print(f"golden {name} {method}: {ex.scene.builder.node_count} nodes", flush=True)
        if name == "cornell":
            assert_tree_matches(a, g, f"{name}_{method}")
        assert gb.digest(a) == str(g[f"{name}_{method}_sha256"])


def _om_vertices(n, seed, kind):
    rng = np.random.default_rng(seed)
    if kind == "random":
        v0 = rng.normal(size=(n, 3)) * 10.0
        return v0, v0 + rng.normal(size=(n, 3)) * 0.3, v0 + rng.normal(size=(n, 3)) * 0.3
    if kind == "grid":
        # coarse grid and repeated triangles: equal centroid coordinates,
        # whose current order a stable sort must keep
        v0 = rng.integers(0, 6, size=(n, 3)).astype(np.float64)
        return v0, v0 + [1.0, 0.0, 0.0], v0 + [0.0, 1.0, 0.0]
    if kind == "flat":
        # z = 0 plane with equal x and y extents: longest-axis ties (x wins)
        v0 = np.stack([rng.integers(0, 16, n), rng.integers(0, 16, n), np.zeros(n)], 1).astype(np.float64)
        return v0, v0 + [1.0, 0.0, 0.0], v0 + [0.0, 1.0, 0.0]
    # degenerate triangles at the origin with either sign of zero: -0.0 and
    # +0.0 compare equal in Python's sort, so their order must stay
    v0 = np.where(rng.random((n, 3)) < 0.5, -0.0, 0.0)
    v0[: n // 4] = rng.normal(size=(n // 4, 3))
    return v0, v0.copy(), v0.copy()


@pytest.mark.parametrize("n,leaf,kind", [(1, 4, "random"), (4, 4, "random"), (5, 4, "random"), (37, 16, "random"),
                                         (3000, 4, "random"), (3000, 4, "grid"), (2000, 1, "flat"),
                                         (500, 4, "zeros"), (60000, 4, "random")])
def test_object_median_build_matches_oracle(n, leaf, kind):
    """core/bvh.py's rule on the GPU: node arrays and leaf order equal the
    oracle's fp64 restatement bit for bit."""
    from roxy_dcc_amd.bvh import build_object_median

    v0, v1, v2 = _om_vertices(n, 7 + n, kind)
    print(f"object median n={n} root leaf={leaf} kind={kind}", flush=True)
    b = build_object_median(v0, v1, v2, max_leaf_size=leaf)
    _assert_tree_equal(b, bvh_oracle.build_object_median(v0, v1, v2, max_leaf=leaf))


def test_object_median_build_matches_reference_golden():
    """The GPU build against the reference's own BVHNode output
    (tests/golden/object_median.npz): preorder ids, leaf ranges and order,
    boxes rounded to fp32."""
    from pathlib import Path

    from roxy_dcc_amd.bvh import build_object_median

    g = np.load(Path(__file__).resolve().parent / "golden" / "object_median.npz")
    for k, (n, leaf, _seed) in enumerate(g["cases"]):
        print(f"golden case {k}: n={n} root leaf={leaf}", flush=True)
        b = build_object_median(g[f"case{k}_v0"], g[f"case{k}_v1"], g[f"case{k}_v2"], max_leaf_size=int(leaf))
        r = b.reference_arrays()
        assert b.node_count == len(g[f"case{k}_left"])
        np.testing.assert_array_equal(b.ordered_tris, g[f"case{k}_order"])
        for f in ("left", "right", "tri_start", "tri_count"):
            np.testing.assert_array_equal(r[f], g[f"case{k}_{f}"], err_msg=f"{k} {f}")
        for f in ("aabb_min", "aabb_max"):
            np.testing.assert_array_equal(r[f], g[f"case{k}_{f}"].astype(np.float32), err_msg=f"{k} {f}")


def test_object_median_scene_renders_like_the_oracle():
    """A scene built with method="object_median" traverses and shades
    exactly as the oracle over its own object-median tree."""
    out = helpers.render_pair("mesh", 32, 32, samples=2, bvh_method="object_median")
    b = out["extract"].scene.builder
    tb = out["extract"].flat.triangles()
    _assert_tree_equal(b, bvh_oracle.build_object_median(tb.v0, tb.v1, tb.v2, 4))
    same = np.mean(np.all(out["gpu_color"] == out["cpu_color"], axis=1))
    print(f"object-median mesh: exact {same:.4f} rays {out['gpu_rays']} {out['cpu_rays']}", flush=True)
    assert out["gpu_rays"] == out["cpu_rays"]
    assert same == 1.0


def test_packed_nodes_match_reference_layout():
    """64-byte nodes hold the children's boxes and links of fields.py's layout."""
    rng = np.random.default_rng(5)
    n = 20000
    v0 = rng.uniform(0, 10, size=(n, 3)).astype(np.float32)
    v1 = (v0 + 0.1).astype(np.float32)
    v2 = (v0 + np.float32([0, 0.1, 0.2])).astype(np.float32)
    b = _gpu_build(v0, v1, v2, 8)
    r = b.reference_arrays()
    raw = b.d["nodes"].cpu().numpy().view(np.uint8).reshape(-1, 64)
    boxes = raw[:, :48].copy().view(np.float32).reshape(-1, 12)
    links = raw[:, 48:64].copy().view(np.int32).reshape(-1, 4)
    hdr = b.d["header"].cpu().numpy().view(np.uint8)
    hf = hdr.copy().view(np.float32)
    hi = hdr.copy().view(np.int32)
    np.testing.assert_array_equal(hf[0:3], r["aabb_min"][0])
    np.testing.assert_array_equal(hf[4:7], r["aabb_max"][0])
    assert hi[3] == 0 and hi[7] == b.depth
    interior = np.nonzero(r["left"] >= 0)[0]
    L, R = r["left"][interior], r["right"][interior]
    np.testing.assert_array_equal(boxes[interior, 0:3], r["aabb_min"][L])
    np.testing.assert_array_equal(boxes[interior, 3:6], r["aabb_max"][L])
    np.testing.assert_array_equal(boxes[interior, 6:9], r["aabb_min"][R])
    np.testing.assert_array_equal(boxes[interior, 9:12], r["aabb_max"][R])
    np.testing.assert_array_equal(links[interior, 2], L)
    np.testing.assert_array_equal(links[interior, 3], R)
    for side, child in ((0, L), (1, R)):
        leaf = r["tri_count"][child] > 0
        want = np.where(leaf, ~((r["tri_start"][child] << 8) | r["tri_count"][child]), child)
        np.testing.assert_array_equal(links[interior, side], want)


def _closest_gpu(scene, rays, counters=True):
    from roxy_dcc_amd import _native as nat

    torch = _torch()
    n = rays.shape[1]
    d_rays = torch.from_numpy(rays).cuda()
    out = {k: torch.empty(n, dtype=torch.float32 if k in ("t", "u", "v") else torch.int32, device="cuda")
           for k in ("t", "id", "kind", "u", "v")}
    ctr = torch.zeros(8, dtype=torch.int64, device="cuda")
    nat.check(nat.lib().rx_closest_hit(C.byref(scene.struct()), d_rays.data_ptr(), n, out["t"].data_ptr(),
                                       out["id"].data_ptr(), out["kind"].data_ptr(), out["u"].data_ptr(),
                                       out["v"].data_ptr(), ctr.data_ptr() if counters else None,
                                       nat.stream_ptr()))
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    c = ctr.cpu().numpy()
    res["counts"] = {"node_pops": int(c[2]), "tri_tests": int(c[3]), "prim_tests": int(c[4])}
    return res


def _check_hits(g, o, label):
    kinds = np.bincount(o["kind"], minlength=3)
    print(f"{label}: {len(o['t'])} rays, misses {kinds[0]} prims {kinds[1]} tris {kinds[2]}; "
          f"counts gpu {g['counts']} cpu {o['counts']}", flush=True)
    np.testing.assert_array_equal(g["kind"], o["kind"])
    np.testing.assert_array_equal(g["id"], o["id"])
    np.testing.assert_array_equal(g["t"].view(np.uint32), o["t"].view(np.uint32))
    hit = o["kind"] == 2
    np.testing.assert_array_equal(g["u"][hit].view(np.uint32), o["u"][hit].view(np.uint32))
    np.testing.assert_array_equal(g["v"][hit].view(np.uint32), o["v"][hit].view(np.uint32))
    assert g["counts"]["node_pops"] == o["counts"]["node_pops"]
    assert g["counts"]["tri_tests"] == o["counts"]["tri_tests"]
    assert g["counts"]["prim_tests"] == o["counts"]["prim_tests"]


@pytest.mark.parametrize("name", ["cornell", "mesh", "large"])
def test_closest_hit_matches_oracle(name):
    """Camera rays and 10^6 random rays (origins in the scene box,
    directions uniform on the sphere) per configuration, the 2M-triangle
    scene included."""
    from roxy_dcc_amd.extract import extract_scene

    world, w, h, _ = helpers.load(name)
    ex = extract_scene(world)
    osc = helpers.flat_to_oracle(ex.flat)
    cam_rays = helpers.camera_rays_np(world.active_camera, 480 if name == "large" else 128,
                                      270 if name == "large" else 128)
    rnd = helpers.random_rays(ex.flat, 1_000_000)
    for label, rays in (("camera", cam_rays), ("random", rnd)):
        _check_hits(_closest_gpu(ex.scene, rays), orc.closest_hit(osc, rays), f"{name}/{label}")


@pytest.mark.parametrize("name", ["mesh", "large"])
def test_closest_hit_secondary_rays_match_oracle(name):
    """Rays leaving hit points (offset 1e-4 along the oriented normal), as the
    bounce loop makes them (trace.py:234)."""
    from roxy_dcc_amd.extract import extract_scene

    world, *_ = helpers.load(name)
    ex = extract_scene(world)
    osc = helpers.flat_to_oracle(ex.flat)
    rays = helpers.camera_rays_np(world.active_camera, 256, 256)
    o = orc.closest_hit(osc, rays)
    hit = o["kind"] != 0
    nrm = orc.hit_normals(osc, rays, o)
    p = rays[:3].T + rays[3:].T * o["t"][:, None]
    org = (p + nrm * np.float32(1e-4))[hit]
    rng = np.random.default_rng(9)
    d = rng.normal(size=org.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = np.where((d * nrm[hit]).sum(1, keepdims=True) < 0, -d, d)
    sec = np.ascontiguousarray(np.concatenate([org.T, d.T]).astype(np.float32))
    _check_hits(_closest_gpu(ex.scene, sec), orc.closest_hit(osc, sec), f"{name}/secondary")


def test_any_hit_matches_oracle():
    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd.extract import extract_scene

    torch = _torch()
    world, *_ = helpers.load("mesh")
    ex = extract_scene(world)
    osc = helpers.flat_to_oracle(ex.flat)
    rays = helpers.random_rays(ex.flat, 100_000, seed=3)
    tmax = np.random.default_rng(4).uniform(0.0, 1.5, size=rays.shape[1]).astype(np.float32)
    want = orc.any_hit(osc, rays, tmax)
    d_rays, d_t = torch.from_numpy(rays).cuda(), torch.from_numpy(tmax).cuda()
    occ = torch.empty(rays.shape[1], dtype=torch.int32, device="cuda")
    nat.check(nat.lib().rx_any_hit(C.byref(ex.scene.struct()), d_rays.data_ptr(), d_t.data_ptr(), rays.shape[1],
                                   occ.data_ptr(), None, nat.stream_ptr()))
    got = occ.cpu().numpy()
    print(f"any-hit occluded fraction {want.mean():.3f}", flush=True)
    np.testing.assert_array_equal(got, want)


# the last two are passes of 2^19+ path slots: the 1024-thread traversal
# blocks with the LDS treelet, pixel-major primaries and sorted secondary
# queues (smaller passes run the 64-thread shape)
@pytest.mark.parametrize("name,size,spp", [("cornell", 48, 8), ("mesh", 40, 4), ("mesh", 512, 2),
                                           ("large", 768, 1)])
def test_render_matches_oracle(name, size, spp):
    res = helpers.render_pair(name, width=size, height=size, samples=spp)
    g, c = res["gpu_color"], res["cpu_color"]
    same = np.all(g == c, axis=1)
    print(f"{name}: exact pixels {same.mean():.4f} max|d| {np.abs(g - c).max():.3g} rays gpu {res['gpu_rays']} "
          f"cpu {res['cpu_rays']} mean {g.mean(0)}", flush=True)
    assert np.isfinite(g).all()
    np.testing.assert_array_equal(g, c)
    np.testing.assert_array_equal(res["gpu_normal"], res["cpu_normal"])
    np.testing.assert_array_equal(res["gpu_albedo"], res["cpu_albedo"])
    np.testing.assert_array_equal(res["gpu_depth"], res["cpu_depth"])
    assert res["gpu_rays"] == res["cpu_rays"]
    assert res["gpu_counts"]["shadow_rays"] == res["cpu_counts"]["shadow_rays"]
    disp = np.sqrt(np.minimum(g, 1.0))
    np.testing.assert_array_equal(res["gpu_display"], disp.astype(np.float32))


@pytest.mark.parametrize("width,height,spp", [(515, 509, 2), (250, 130, 4)])
def test_untiled_frame_pixel_major_matches_oracle(width, height, spp):
    """Frames whose sides are not multiples of 8 keep row-major path slots;
    their camera rays are still traced pixel-major, computed by the
    traversal (at 515x509: a pass of 2^19+ slots, on the eye-relative copies,
    with primitive masks over the partial tiles at the right and bottom
    edges)."""
    res = helpers.render_pair("mesh", width=width, height=height, samples=spp)
    g, c = res["gpu_color"], res["cpu_color"]
    print(f"mesh {width}x{height}x{spp}: exact pixels {np.all(g == c, axis=1).mean():.4f} rays gpu "
          f"{res['gpu_rays']} cpu {res['cpu_rays']}", flush=True)
    np.testing.assert_array_equal(g, c)
    np.testing.assert_array_equal(res["gpu_normal"], res["cpu_normal"])
    np.testing.assert_array_equal(res["gpu_depth"], res["cpu_depth"])
    assert res["gpu_rays"] == res["cpu_rays"]
    assert res["gpu_counts"]["shadow_rays"] == res["cpu_counts"]["shadow_rays"]


@pytest.mark.parametrize("name,shape", [("large", (1920, 1080, 16)), ("mesh", (512, 512, 64)),
                                        ("cornell", (256, 256, 16))])
def test_benchmarked_render_matches_oracle_on_sampled_pixels(name, shape):
    """The benchmarked workloads themselves at their full sizes, each in one
    pass as bench.py renders them (the 2M-triangle 1080p scene on the
    link-free tree copy, eye-relative camera rays, tile masks and sorted
    queues): every buffer bit of 4096 random pixels equals the fp32
    oracle's render of those pixels."""
    import torch

    from oracle import oracle as orc
    from roxy_dcc_amd import renderer as R
    from roxy_dcc_amd.extract import extract_scene

    world, W, H, st = helpers.load(name)
    spp = st["samples"]
    assert (W, H, spp) == shape
    ex = extract_scene(world)
    integ = R.Integrator(ex.scene, W, H, spp)
    frame = R.DeviceFrame(W, H)
    cam = world.active_camera
    mode = R.TaichiRenderer._direct_light_mode_to_id(st["direct_light_mode"])
    print(f"{name} at {W}x{H}x{spp}: one GPU pass, oracle on 4096 pixels", flush=True)
    prm = R.params_struct(st["max_depth"], int(world._use_sky), mode, st["direct_light_max_depth"], 10.0,
                          world._background_color, 0, spp)
    integ.render(frame, R.camera_struct(cam, W, H), prm, 0, spp)
    torch.cuda.synchronize()
    pix = np.random.default_rng(7).choice(W * H, size=4096, replace=False).astype(np.int32)
    acc = orc.Accum(len(pix))
    oprm = orc.params(st["max_depth"], int(world._use_sky), mode, st["direct_light_max_depth"], 10.0,
                      world._background_color, 0)
    orc.render(helpers.flat_to_oracle(ex.flat), helpers.oracle_camera(cam, W, H), oprm, 0, spp, acc, pixels=pix)
    np.testing.assert_array_equal(frame.color.cpu().numpy().reshape(-1, 3)[pix], acc.color)
    np.testing.assert_array_equal(frame.normal.cpu().numpy().reshape(-1, 3)[pix], acc.normal)
    np.testing.assert_array_equal(frame.albedo.cpu().numpy().reshape(-1, 3)[pix], acc.albedo)
    np.testing.assert_array_equal(frame.depth.cpu().numpy().reshape(-1)[pix], acc.depth)


def test_render_independent_of_pass_split():
    a = helpers.render_pair("cornell", width=32, height=32, samples=6, spp_per_pass=6)
    b = helpers.render_pair("cornell", width=32, height=32, samples=6, spp_per_pass=1)
    np.testing.assert_array_equal(a["gpu_color"], b["gpu_color"])
    np.testing.assert_array_equal(a["gpu_depth"], b["gpu_depth"])


def _render_mesh(flags, W=128, H=128, spp=4):
    import torch

    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd import renderer as R
    from roxy_dcc_amd.extract import extract_scene

    world, _, _, st = helpers.load("mesh")
    ex = extract_scene(world)
    integ = R.Integrator(ex.scene, W, H, spp)
    frame = R.DeviceFrame(W, H)
    prm = R.params_struct(st["max_depth"], int(world._use_sky), 0, st["direct_light_max_depth"], 10.0,
                          world._background_color, 3, spp, flags | nat.RX_FLAG_PROFILE | nat.RX_FLAG_COUNT_TRAFFIC)
    integ.render(frame, R.camera_struct(world.active_camera, W, H), prm, 0, spp)
    torch.cuda.synchronize()
    return frame.color.cpu().numpy(), integ.counts(), integ.profile


def test_reordered_secondary_queues_change_nothing():
    """k_extend traces each secondary queue sorted by direction and origin
    and writes every hit back to its ray's slot: the image and every
    traversal counter equal those of the queue traced in shade's order
    (primitive tests aside: camera rays test the primitives of the tiles of
    the rays a wave takes together, which the primary queue's order sets)."""
    from roxy_dcc_amd import _native as nat

    # 512 x 512 x 4 = 2^20 path slots: large enough to sort
    sorted_img, sorted_cnt, sorted_prof = _render_mesh(0, 512, 512, 4)
    plain_img, plain_cnt, plain_prof = _render_mesh(nat.RX_FLAG_KEEP_QUEUE_ORDER, 512, 512, 4)
    print(f"reorder launches {sorted_prof.n_reorder} ({sorted_prof.ms_reorder:.2f} ms); extend "
          f"{sorted_prof.ms_extend:.2f} vs {plain_prof.ms_extend:.2f} ms", flush=True)
    assert sorted_prof.n_reorder >= 3 and plain_prof.n_reorder == 0
    np.testing.assert_array_equal(sorted_img, plain_img)
    assert {k: v for k, v in sorted_cnt.items() if k != "prim_tests"} == \
        {k: v for k, v in plain_cnt.items() if k != "prim_tests"}


def test_render_all_lights_mode_matches_oracle():
    from roxy_dcc_amd import scenes

    world_tuple = scenes.large_scene(seed=3, n_meshes=20, seg_u=20, seg_v=11, n_spheres=16)
    res = helpers.render_pair("large", width=64, height=36, samples=2, direct_light_mode="all",
                              world_tuple=world_tuple)
    same = np.all(res["gpu_color"] == res["cpu_color"], axis=1)
    print(f"large/all: exact {same.mean():.4f} shadow rays {res['gpu_counts']['shadow_rays']}", flush=True)
    np.testing.assert_array_equal(res["gpu_color"], res["cpu_color"])
    assert res["gpu_rays"] == res["cpu_rays"]
    assert res["gpu_counts"]["shadow_rays"] == res["cpu_counts"]["shadow_rays"]


def test_render_prim_prepass_matches_oracle():
    """Closest-primitive scan as its own pass over each ray queue (forced by
    RX_FLAG_PRIM_PREPASS; 16 analytic spheres, 8x8 tiled slot order)."""
    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd import scenes

    world_tuple = scenes.large_scene(seed=4, n_meshes=20, seg_u=20, seg_v=11, n_spheres=16)
    res = helpers.render_pair("large", width=64, height=40, samples=2, world_tuple=world_tuple,
                              flags=nat.RX_FLAG_PRIM_PREPASS | nat.RX_FLAG_COUNT_TRAFFIC)
    same = np.all(res["gpu_color"] == res["cpu_color"], axis=1)
    g, c = res["gpu_counts"], res["cpu_counts"]
    print(f"large/prepass: exact {same.mean():.4f} rays {g['rays']} {c['rays']} prim tests {g['prim_tests']} "
          f"{c['prim_tests']}", flush=True)
    np.testing.assert_array_equal(res["gpu_color"], res["cpu_color"])
    # same paths, so the same traversal work, pop for pop
    for k in ("rays", "node_pops", "tri_tests", "prim_tests", "shadow_rays", "shadow_prim_tests"):
        assert g[k] == c[k], k


@pytest.mark.parametrize("width,height", [(96, 64), (384, 216)])
def test_render_prim_prepass_pixel_major_matches_oracle(width, height):
    """The pre-pass in passes whose camera rays the traversal computes
    pixel-major (a BVH of at least 16384 triangles, two samples per pass):
    each camera ray reads the pre-pass's result at its sample-major hit
    slot.  At 384x216 the camera rays also run on the eye-relative copies."""
    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd import scenes

    world_tuple = scenes.large_scene(seed=6, n_meshes=40, seg_u=20, seg_v=11, n_spheres=16)
    print(f"large/prepass pixel-major at {width}x{height}", flush=True)
    res = helpers.render_pair("large", width=width, height=height, samples=4, spp_per_pass=2,
                              world_tuple=world_tuple, flags=nat.RX_FLAG_PRIM_PREPASS | nat.RX_FLAG_COUNT_TRAFFIC)
    np.testing.assert_array_equal(res["gpu_color"], res["cpu_color"])
    np.testing.assert_array_equal(res["gpu_depth"], res["cpu_depth"])
    g, c = res["gpu_counts"], res["cpu_counts"]
    for k in ("rays", "node_pops", "tri_tests", "prim_tests", "shadow_rays", "shadow_prim_tests"):
        assert g[k] == c[k], k


@pytest.mark.parametrize("radius,dilation,guides", [(1, 1, "none"), (1, 1, "all"), (2, 1, "normal"),
                                                    (1, 2, "all"), (1, 4, "albedo"), (1, 16, "depth"),
                                                    (3, 3, "all")])
def test_denoise_matches_oracle(radius, dilation, guides):
    from roxy_dcc_amd.denoise import edge_aware_denoise

    rng = np.random.default_rng(radius * 100 + dilation)
    h, w = 37, 53
    img = rng.random((h, w, 3), dtype=np.float32) * 2
    kw = {}
    if guides in ("all", "normal"):
        kw["normal"] = rng.normal(size=(h, w, 3)).astype(np.float32)
    if guides in ("all", "albedo"):
        kw["albedo"] = rng.random((h, w, 3), dtype=np.float32)
    if guides in ("all", "depth"):
        kw["depth"] = rng.random((h, w), dtype=np.float32) * 5
    want = denoise_oracle.denoise(img, radius, 0.3, 0.8, dilation=dilation, **kw)
    got = edge_aware_denoise(img, radius, 0.3, 0.8, dilation=dilation, **kw)
    err = np.abs(got - want).max()
    print(f"denoise r{radius} d{dilation} {guides}: max err {err:.3g}", flush=True)
    assert err <= 1e-5


def test_denoise_matches_reference_golden():
    """rx_denoise against the reference's own edge_aware_denoise output
    (tests/golden/denoise.npz): every guide combination, radii 1-3, depth as
    HxW and HxWx1, clipped amount and the identity cases, within 1e-5."""
    from pathlib import Path

    from roxy_dcc_amd.denoise import edge_aware_denoise
    from tests.test_reference_goldens import denoise_inputs

    g = np.load(Path(__file__).resolve().parent / "golden" / "denoise.npz")
    worst = 0.0
    for c in range(len(g["cases"])):
        color, radius, sc, amount, kw = denoise_inputs(g, c)
        got = np.asarray(edge_aware_denoise(color, radius, sc, amount, **kw), np.float32)
        err = float(np.abs(got - g[f"case{c}_out"]).max())
        worst = max(worst, err)
        assert err <= 1e-5, (c, err)
    print(f"denoise vs reference golden: max |diff| {worst:.3g} over {len(g['cases'])} cases", flush=True)
    from roxy_dcc_amd.denoise import linear_to_display

    np.testing.assert_array_equal(linear_to_display(g["display_in"]), g["display_out"])


def test_display_matches_reference_values():
    from roxy_dcc_amd.denoise import linear_to_display

    out = linear_to_display(np.array([[[-1.0, 0.25, 4.0]]], dtype=np.float32))
    assert out[0, 0, 0] == 0.0 and out[0, 0, 1] == 0.5 and out[0, 0, 2] == 1.0


def test_taichi_renderer_end_to_end():
    from roxy_dcc_amd import TaichiRenderer
    from roxy_dcc_amd.world import Image

    world, *_ = helpers.load("cornell")
    img = Image(64, 64)
    r = TaichiRenderer(world, img, None, samples=8, max_depth=5, denoise=True, verbose=False)
    r.render()
    assert np.isfinite(img.pixels).all() and img.pixels.max() <= 1.0 and img.pixels.min() >= 0.0
    s = r.last_stats
    assert s.samples_rendered == 8 and s.bvh_triangles == 36 and s.bvh_nodes == 15
    assert r.last_ray_count == s.rays_cast > 64 * 64 * 8
    # the image holds the device's display buffer bit for bit, denoised or not
    assert np.array_equal(img.pixels, r._denoised_display().cpu().numpy())
    plain = Image(48, 32)
    r = TaichiRenderer(world, plain, None, samples=4, max_depth=5, denoise=False, verbose=False)
    r.render()
    assert np.array_equal(plain.pixels, r.frame.display.cpu().numpy())
    assert plain.pixels.max() > 0.0


@pytest.mark.parametrize("interval", [1, 3])
def test_adaptive_stop_matches_oracle(interval):
    """The adaptive stop of rendering/ray_tracer.py:531-541 on the device:
    the run stops at the same sample as the rule applied to the oracle's
    per-sample errors (a threshold with a clear margin), mid-pass, and its
    image, ray count and error equal the oracle's after that many samples."""
    from oracle import oracle as orc
    from roxy_dcc_amd import renderer as R
    from roxy_dcc_amd.world import Image

    from roxy_dcc_amd.extract import flatten_scene

    world, _, _, st = helpers.load("cornell")      # square camera, rendered at 32x32
    W = H = 32
    samples, amin = 48, 4

    osc = helpers.flat_to_oracle(flatten_scene(world))
    ocam = helpers.oracle_camera(world.active_camera, W, H)
    oprm = orc.params(st["max_depth"], int(world._use_sky), 0, st["direct_light_max_depth"], 10.0,
                      world._background_color, 9)
    errs = orc.adaptive_errors(osc, ocam, oprm, samples, orc.Accum(W * H))
    target = thr = None
    for n in range(amin + 5, samples - 4):
        if (n - amin) % interval:
            continue
        earlier = [errs[m] for m in range(amin, n) if (m - amin) % interval == 0]
        if errs[n] < 0.97 * min(earlier):
            target, thr = n, 0.5 * (errs[n] + min(earlier))
            break
    assert target is not None
    assert orc.adaptive_stop(errs, samples, amin, thr, interval) == target
    acc = orc.Accum(W * H)
    orc.render(osc, ocam, oprm, 0, target, acc)

    img = Image(W, H)
    r = R.TaichiRenderer(world, img, None, samples=samples, max_depth=st["max_depth"], seed=9, samples_per_pass=8,
                         adaptive_sampling=True, adaptive_min_samples=amin, adaptive_threshold=thr,
                         adaptive_check_interval=interval, verbose=False)
    r.render()
    stats = r.last_stats
    got = r.frame.color.cpu().numpy().reshape(-1, 3)
    same = np.all(got == acc.color, axis=1).mean()
    print(f"adaptive interval {interval}: stop {stats.samples_rendered} (oracle {target}), error "
          f"{stats.adaptive_error:.6f} (oracle {errs[target]:.6f}), exact {same:.4f}, rays {r.last_ray_count} "
          f"({acc.counts.rays})", flush=True)
    assert stats.adaptive_stopped and stats.samples_rendered == target
    np.testing.assert_array_equal(got, acc.color)
    assert r.last_ray_count == acc.counts.rays
    assert abs(stats.adaptive_error - errs[target]) <= 1e-5 * max(errs[target], 1e-3)


@pytest.mark.parametrize("flip_v", [True, False])
def test_textured_render_matches_oracle(flip_v):
    """Textured albedo (trace.py:96-136) in the wavefront shade stage: the
    render and its first-hit albedo guide equal the oracle's."""
    world, tex, _ = helpers.textured_floor_world(16, 8, flip_v=flip_v)
    settings = dict(max_depth=3, direct_light_mode="one", direct_light_max_depth=1)
    res = helpers.render_pair("floor", samples=3, world_tuple=(world, 48, 48, settings))
    same = np.all(res["gpu_color"] == res["cpu_color"], axis=1).mean()
    alb = np.all(res["gpu_albedo"] == res["cpu_albedo"], axis=1).mean()
    print(f"textured floor flip_v={flip_v}: exact colour {same:.4f}, exact albedo {alb:.4f}", flush=True)
    np.testing.assert_array_equal(res["gpu_color"], res["cpu_color"])
    np.testing.assert_array_equal(res["gpu_albedo"], res["cpu_albedo"])
    assert res["gpu_rays"] == res["cpu_rays"]


def test_obj_asset_render_matches_oracle(tmp_path):
    """OBJ/MTL on disk -> OBJReader -> extraction -> render: equal to the
    oracle's render of the same flattened scene (a box room, a tilted
    quad per material: diffuse, metal, glossy, emissive)."""
    from roxy_dcc_amd.assets import OBJReader
    from roxy_dcc_amd.world import Camera, World

    (tmp_path / "room.mtl").write_text("\n".join([
        "newmtl white", "Kd 0.73 0.73 0.73", "newmtl red", "Kd 0.65 0.05 0.05",
        "newmtl mirror", "illum 3", "Kd 0.9 0.9 0.9", "Ns 800",
        "newmtl satin", "Ks 0.4 0.4 0.4", "Kd 0.2 0.5 0.7", "Ns 300",
        "newmtl lamp", "Ke 12 12 12", "Kd 1 1 1"]))
    v = ["v 0 0 0", "v 1 0 0", "v 1 0 1", "v 0 0 1", "v 0 1 0", "v 1 1 0", "v 1 1 1", "v 0 1 1",
         "v 0.35 0.999 0.35", "v 0.65 0.999 0.35", "v 0.65 0.999 0.65", "v 0.35 0.999 0.65",
         "v 0.2 0.1 0.3", "v 0.6 0.1 0.2", "v 0.6 0.5 0.5", "v 0.2 0.5 0.6"]
    f = ["usemtl white", "f 1 2 3 4", "f 5 8 7 6", "f 4 3 7 8", "usemtl red", "f 1 4 8 5",
         "usemtl mirror", "f 2 6 7 3", "usemtl satin", "f 13 14 15 16", "usemtl lamp", "f 9 12 11 10"]
    (tmp_path / "room.obj").write_text("\n".join(["mtllib room.mtl"] + v + ["o room"] + f) + "\n")
    w = World(background_color=(0.0, 0.0, 0.0), use_sky=False)
    w.add_object(OBJReader.load_indexed(tmp_path / "room.obj"))
    w.add_camera(Camera(position=(0.5, 0.5, -1.3), forward=(0, 0, 1), up=(0, 1, 0), fov=40, width=40, height=40))
    settings = dict(max_depth=5, direct_light_mode="one", direct_light_max_depth=1)
    res = helpers.render_pair("room", samples=4, world_tuple=(w, 40, 40, settings))
    same = np.all(res["gpu_color"] == res["cpu_color"], axis=1).mean()
    print(f"OBJ room: {res['extract']['bvh_triangles']} triangles, {res['extract']['bvh_materials']} materials, "
          f"exact {same:.4f}", flush=True)
    assert res["extract"]["bvh_triangles"] == 14 and res["extract"]["bvh_materials"] == 5
    np.testing.assert_array_equal(res["gpu_color"], res["cpu_color"])
    assert res["gpu_rays"] == res["cpu_rays"]
