"""Pins for the CPU oracle (no GPU needed).

The oracle is only trusted after it matches what the reference's own tests
hold and what the mathematics fixes:
  * Philox4x32-10 known-answer vectors (Random123 kat_vectors)
  * sin/cos against libm, jitter-hash sine against libm
  * known answers of tests/mesh_tests.py:13-43,105-112 (Moeller-Trumbore),
    tests/shapes_tests.py:53-105 (sphere), tests/hit_record_tests.py:40-61
    (orientation), tests/world_tests.py:164-187 (sky)
  * brute-force closest hit over every triangle (pins BVH build +
    traversal: a missed closer triangle fails it)
  * ray counts of tests/ray_tracer_tests.py:107-122 (empty scene: one ray
    per pixel-sample; depth 0: none)
  * closed-form renders: a diffuse or metal floor under a constant
    background returns its albedo exactly; a floor under one sphere light
    has expectation 2 a Le sin^2(theta_max) (next-event estimate plus the
    reference's implicit hit of the same light)
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import bvh_oracle
from oracle import oracle as orc
from tests import helpers


# ------------------------------------------------------------------ Philox
@pytest.mark.parametrize("ctr,key,want", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, want):
    assert orc.philox(ctr, key) == list(want)


def test_sincos_turns_accuracy():
    u = np.random.default_rng(0).random(1_000_000, dtype=np.float32)
    u = np.concatenate([u, np.float32([0.0, 0.125, 0.25, 0.5, 0.75, 0.999999])])
    s, c = orc.sincos_turns(u)
    ph = 2.0 * np.pi * u.astype(np.float64)
    assert np.abs(s - np.sin(ph)).max() < 3e-7
    assert np.abs(c - np.cos(ph)).max() < 3e-7
    assert s[-6] == 0.0 and c[-6] == 1.0


def test_hash_sine_matches_libm():
    x = np.concatenate([np.random.default_rng(1).uniform(-1e6, 1e6, 100_000), [0.0, 1.0, 127.1, 580000.3]])
    got = orc.hash_sine(x)
    assert np.abs(got - np.sin(x)).max() < 1e-12


# ------------------------------------------------------------------ scenes
def _tri_scene(tris, prims=None, mats=None):
    tris = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    n = len(tris)
    nrm = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    nrm = nrm.astype(np.float32)
    palette = mats or [(0, (0.5, 0.5, 0.5), 0.0, 1.0, 0.0, -1)]
    return orc.OracleScene(tris[:, 0], tris[:, 1], tris[:, 2], nrm, nrm, nrm, np.zeros(n, np.int32), palette,
                           prims)


def _rays(o, d):
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    return np.ascontiguousarray(np.concatenate([o.T, d.T]))


UNIT_TRI = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]


def test_triangle_known_answers():
    sc = _tri_scene([UNIT_TRI])
    h = orc.closest_hit(sc, _rays([[0.25, 0.25, -1], [2, 2, -1], [0.25, 0.25, 0], [0.25, 0.25, 2], [0.3, 0.2, -1]],
                                  [[0, 0, 1], [0, 0, 1], [1, 0, 0], [0, 0, 1], [0, 0, 1]]))
    assert h["kind"].tolist() == [2, 0, 0, 0, 2]
    assert h["t"][0] == 1.0
    assert h["t"][1] >= 1e9 and h["t"][2] >= 1e9 and h["t"][3] >= 1e9
    u, v = h["u"][4], h["v"][4]
    assert u >= 0 and v >= 0 and u + v <= 1


def test_closest_of_two_triangles():
    sc = _tri_scene([[[0, 0, 1], [1, 0, 1], [0, 1, 1]], [[0, 0, 3], [1, 0, 3], [0, 1, 3]]])
    h = orc.closest_hit(sc, _rays([[0.25, 0.25, -1]], [[0, 0, 1]]))
    assert abs(h["t"][0] - 2.0) < 1e-6


def _sphere_prims(center, radius, mat_type=0, emission=0.0, albedo=(0.5, 0.5, 0.5)):
    c = np.asarray(center, np.float32).reshape(-1, 3)
    k = len(c)
    return {"obj_type": np.zeros(k, np.int32), "obj_center": c, "obj_radius": np.full(k, radius, np.float32),
            "obj_mat_type": np.full(k, mat_type, np.int32), "obj_albedo": np.tile(np.float32(albedo), (k, 1)),
            "obj_emission": np.full(k, emission, np.float32), "obj_ior": np.ones(k, np.float32),
            "obj_roughness": np.zeros(k, np.float32)}


def test_sphere_known_answers():
    sc = orc.OracleScene(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)),
                         np.zeros((0, 3)), np.zeros(0, np.int32), [], _sphere_prims([0, 0, 0], 1.0))
    rays = _rays([[0, 0, -5], [0, 0, 0], [0, 5, -5], [0, 0, 5]], [[0, 0, 1]] * 4)
    h = orc.closest_hit(sc, rays)
    assert h["kind"].tolist() == [1, 1, 0, 0]
    assert abs(h["t"][0] - 4.0) < 1e-6 and abs(h["t"][1] - 1.0) < 1e-6
    n = orc.hit_normals(sc, rays, h)
    # oriented against the ray (trace.py:161-163; hit_record_tests.py:40-61)
    np.testing.assert_allclose(n[0], [0, 0, -1], atol=1e-6)
    np.testing.assert_allclose(n[1], [0, 0, -1], atol=1e-6)


def test_plane_and_cube_known_answers():
    prims = {"obj_type": np.int32([1, 2]), "obj_normal": np.float32([[0, 1, 0], [0, 0, 0]]),
             "obj_offset": np.float32([0, 0]), "obj_center": np.float32([[0, 0, 0], [5, 5, 5]]),
             "obj_extra": np.float32([[0, 0, 0], [0.5, 0.5, 0.5]]), "obj_mat_type": np.int32([0, 0])}
    sc = orc.OracleScene(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)),
                         np.zeros((0, 3)), np.zeros(0, np.int32), [], prims)
    rays = _rays([[0, 1, 0], [3, 5, 5], [0, 1, 0]], [[0, -1, 0], [1, 0, 0], [1, 0, 0]])
    h = orc.closest_hit(sc, rays)
    assert h["kind"].tolist() == [1, 1, 0]
    assert h["id"][0] == 0 and h["t"][0] == 1.0
    assert h["id"][1] == 1 and h["t"][1] == 1.5
    n = orc.hit_normals(sc, rays, h)
    np.testing.assert_array_equal(n[0], [0, 1, 0])
    np.testing.assert_array_equal(n[1], [-1, 0, 0])


def test_primitive_beats_bvh_on_ties():
    """A sphere and a triangle at the same t: the primitive scan wins
    (intersect.py:292 accepts the BVH only if strictly closer)."""
    tri = [[-1, -1, 1], [1, -1, 1], [0, 1, 1]]
    sc = orc.OracleScene(*[np.float32([p]) for p in tri], *[np.float32([[0, 0, -1]])] * 3, np.zeros(1, np.int32),
                         [(0, (0.5, 0.5, 0.5), 0.0, 1.0, 0.0, -1)], _sphere_prims([0, 0, 2], 1.0))
    h = orc.closest_hit(sc, _rays([[0, 0, 0]], [[0, 0, 1]]))
    assert h["kind"][0] == 1 and h["t"][0] == 1.0


def _brute_force(tris, rays):
    """fp64 plane-hit + barycentric inside test over every triangle."""
    o = rays[:3].T.astype(np.float64)
    d = rays[3:].T.astype(np.float64)
    a, b, c = (tris[:, i].astype(np.float64) for i in range(3))
    nrm = np.cross(b - a, c - a)
    best_t = np.full(len(o), np.inf)
    second = np.full(len(o), np.inf)
    best_i = np.full(len(o), -1)
    for k in range(len(tris)):
        den = d @ nrm[k]
        with np.errstate(divide="ignore", invalid="ignore"):
            t = ((a[k] - o) @ nrm[k]) / den
        p = o + d * t[:, None]
        e0 = np.cross(b[k] - a[k], p - a[k]) @ nrm[k]
        e1 = np.cross(c[k] - b[k], p - b[k]) @ nrm[k]
        e2 = np.cross(a[k] - c[k], p - c[k]) @ nrm[k]
        inside = (e0 >= 0) & (e1 >= 0) & (e2 >= 0) & (np.abs(den) > 1e-12) & (t > 0.001)
        t = np.where(inside, t, np.inf)
        closer = t < best_t
        second = np.where(closer, best_t, np.minimum(second, t))
        best_i = np.where(closer, k, best_i)
        best_t = np.where(closer, t, best_t)
    return best_t, best_i, second


@pytest.mark.parametrize("builder", ["morton", "median", "object_median"])
def test_closest_hit_matches_brute_force(builder):
    rng = np.random.default_rng(12)
    n = 600
    base = rng.uniform(0, 10, size=(n, 3))
    tris = np.stack([base, base + rng.normal(size=(n, 3)), base + rng.normal(size=(n, 3))], axis=1).astype(np.float32)
    nrm = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]).astype(np.float32)
    sc = orc.OracleScene(tris[:, 0], tris[:, 1], tris[:, 2], nrm, nrm, nrm, np.zeros(n, np.int32),
                         [(0, (1, 1, 1), 0, 1, 0, -1)], builder=builder, max_leaf=4)
    m = 4000
    o = rng.uniform(-2, 12, size=(m, 3))
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.ascontiguousarray(np.concatenate([o.T, d.T]).astype(np.float32))
    h = orc.closest_hit(sc, rays)
    bt, bi, second = _brute_force(tris, rays)
    hit_bf = np.isfinite(bt)
    # clear-cut cases: away from triangle edges both agree on hit/miss
    got_hit = h["kind"] == 2
    margin_ok = np.ones(m, bool)
    mismatch = got_hit != hit_bf
    assert mismatch.mean() < 0.002, mismatch.sum()
    both = got_hit & hit_bf & margin_ok
    np.testing.assert_allclose(h["t"][both], bt[both], rtol=1e-4)
    sep = both & (second > bt * (1 + 1e-4))
    orig = sc.tree.order[h["id"][sep]]
    assert (orig == bi[sep]).mean() > 0.999


# ------------------------------------------------------------------ renders
def _camera(pos, fwd, up=(0, 1, 0), fov_deg=40.0, w=8, h=8):
    f = np.asarray(fwd, np.float64)
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64) / np.linalg.norm(up))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return orc.camera(pos, f, r, u, math.tan(math.radians(fov_deg) / 2), w / h, w, h)


EMPTY = dict(v0=np.zeros((0, 3)), v1=np.zeros((0, 3)), v2=np.zeros((0, 3)), n0=np.zeros((0, 3)),
             n1=np.zeros((0, 3)), n2=np.zeros((0, 3)), tri_mat=np.zeros(0, np.int32), palette=[])


@pytest.mark.parametrize("fwd,up,want", [((0, 1, 0), (0, 0, 1), (0.5, 0.7, 1.0)),
                                         ((0, -1, 0), (0, 0, 1), (1.0, 1.0, 1.0)),
                                         ((1, 0, 0), (0, 1, 0), (0.75, 0.85, 1.0))])
def test_sky_known_answers(fwd, up, want):
    sc = orc.OracleScene(**EMPTY)
    cam = _camera((0, 0, 0), fwd, up, fov_deg=1e-4, w=2, h=2)
    acc = orc.Accum(4)
    orc.render(sc, cam, orc.params(max_depth=3, use_sky=1), 0, 3, acc)
    np.testing.assert_allclose(acc.color, np.tile(want, (4, 1)), atol=2e-6)


def test_ray_counts_empty_scene_and_depth_zero():
    sc = orc.OracleScene(**EMPTY)
    cam = _camera((0, 0, 0), (0, 0, 1), w=4, h=4)
    acc = orc.Accum(16)
    orc.render(sc, cam, orc.params(max_depth=4), 0, 3, acc)
    assert acc.counts.rays == 4 * 4 * 3
    acc0 = orc.Accum(16)
    orc.render(sc, cam, orc.params(max_depth=0), 0, 3, acc0)
    assert acc0.counts.rays == 0 and (acc0.color == 0).all()


FLOOR = [[[-100, 0, -100], [100, 0, -100], [100, 0, 100]], [[-100, 0, -100], [100, 0, 100], [-100, 0, 100]]]


@pytest.mark.parametrize("mat", [(0, (0.3, 0.6, 0.9), 0.0, 1.0, 0.0, -1), (1, (0.8, 0.7, 0.2), 0.0, 1.0, 0.0, -1)])
def test_floor_under_constant_background_returns_albedo(mat):
    sc = _tri_scene(FLOOR, mats=[mat])
    cam = _camera((0, 5, 0), (0.3, -1, 0.2), w=8, h=8)
    acc = orc.Accum(64)
    orc.render(sc, cam, orc.params(max_depth=3, bg=(1, 1, 1)), 0, 4, acc)
    np.testing.assert_array_equal(acc.color, np.tile(np.float32(mat[1]), (64, 1)))


@pytest.mark.parametrize("mode", [0, 1])
def test_sphere_light_over_floor_closed_form(mode):
    """E[pixel] = 2 a Le sin^2(theta_max): next-event estimate a Le (1-c^2)
    plus the implicit hit of the light with probability sin^2(theta_max)."""
    a, le, r, height = 0.5, 4.0, 0.5, 3.0
    prims = _sphere_prims([0, height, 0], r, mat_type=3, emission=le, albedo=(1, 1, 1))
    sc = _tri_scene(FLOOR, prims=prims, mats=[(0, (a, a, a), 0.0, 1.0, 0.0, -1)])
    # look straight down from just below the light at a tiny patch
    cam = _camera((0, 1.0, 0), (0, -1, 0), up=(0, 0, 1), fov_deg=1e-3, w=32, h=32)
    acc = orc.Accum(32 * 32)
    spp = 64
    orc.render(sc, cam, orc.params(max_depth=2, direct_light_mode=mode, sample_clamp=0.0, seed=5), 0, spp, acc)
    want = 2 * a * le * (r * r) / (height * height)
    # per-sample variance bound: values in [0, a*le + a*le*2(1-c)]
    sd = (a * le) / math.sqrt(32 * 32 * spp)
    got = acc.color[:, 0].mean()
    assert abs(got - want) < 5 * sd, (got, want, sd)


def test_fp64_oracle_agrees_with_fp32_in_expectation():
    sc = _tri_scene(FLOOR, prims=_sphere_prims([0, 3, 0], 0.5, mat_type=3, emission=4.0, albedo=(1, 1, 1)),
                    mats=[(0, (0.5, 0.5, 0.5), 0.0, 1.0, 0.0, -1)])
    cam = _camera((0, 1.5, -2), (0, -0.5, 1), w=16, h=16)
    a32, a64 = orc.Accum(256), orc.Accum(256, "f64")
    p = orc.params(max_depth=3, seed=2)
    orc.render(sc, cam, p, 0, 64, a32)
    orc.render(sc, cam, p, 0, 64, a64)
    assert np.abs(a32.color.mean() - a64.color.mean()) < 0.01 * max(a64.color.mean(), 1e-3)


def test_second_moments_are_the_squares_of_the_samples():
    """The fp64 instance's per-pixel second moments (the RMSE fixtures'
    per-pixel variance) against the samples themselves: each sample rendered
    alone gives sq = x^2 (x >= 0 after the clamp), and those samples must
    reproduce both the running mean and the summed squares of one call."""
    sc = _tri_scene(FLOOR, prims=_sphere_prims([0, 3, 0], 0.5, mat_type=3, emission=4.0, albedo=(1, 1, 1)),
                    mats=[(0, (0.5, 0.5, 0.5), 0.0, 1.0, 0.0, -1)])
    cam = _camera((0, 1.5, -2), (0, -0.5, 1), w=16, h=16)
    p = orc.params(max_depth=4, seed=3)
    n = 12
    one = orc.Accum(256, "f64", moments=True)
    orc.render(sc, cam, p, 0, n, one)
    xs = []
    for k in range(n):
        a = orc.Accum(256, "f64", moments=True)
        orc.render(sc, cam, p, k, 1, a)
        xs.append(np.sqrt(a.sq))
    xs = np.stack(xs)
    np.testing.assert_allclose(xs.mean(0), one.color, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose((xs ** 2).sum(0), one.sq, rtol=1e-12, atol=1e-12)
    assert (xs.std(0) > 0).any()


def test_render_pixel_subset_matches_full_frame():
    sc = _tri_scene(FLOOR, prims=_sphere_prims([0, 3, 0], 0.5, mat_type=3, emission=4.0, albedo=(1, 1, 1)))
    cam = _camera((0, 1.5, -2), (0, -0.5, 1), w=16, h=16)
    full, part = orc.Accum(256), orc.Accum(10)
    pix = np.arange(0, 256, 26)[:10]
    orc.render(sc, cam, orc.params(max_depth=4, seed=9), 0, 5, full)
    orc.render(sc, cam, orc.params(max_depth=4, seed=9), 0, 5, part, pixels=pix)
    np.testing.assert_array_equal(part.color, full.color[pix])


def test_camera_rays_are_unit_and_jittered_within_pixel():
    cam = _camera((0, 0, 0), (0, 0, 1), w=16, h=8)
    d0 = orc.camera_rays(cam, 0)
    d1 = orc.camera_rays(cam, 1)
    np.testing.assert_allclose(np.linalg.norm(d0, axis=1), 1.0, atol=1e-6)
    assert not np.array_equal(d0, d1)
    # each ray passes through its own pixel: u = (x + jitter)/W*2 - 1 with
    # jitter in [0, 1), and screen-right is fwd x up = -x here
    fov_tan = math.tan(math.radians(20))
    u = (-d0[:, 0] / d0[:, 2]) / (2.0 * fov_tan)
    px = (u + 1) / 2 * 16
    cols = np.tile(np.arange(16), 8)
    assert np.all((px >= cols - 1e-4) & (px < cols + 1 + 1e-4))


# ------------------------------------------------------------------ BVH oracle
def _row_tris(xs):
    v0 = np.float32([[x, 0, 0] for x in xs])
    return v0, v0 + np.float32([1, 0, 0]), v0 + np.float32([0, 1, 0])


def test_bvh_reference_known_answers():
    # tests/bvh_builder_tests.py:20-44,70-75
    assert bvh_oracle.build_morton(*_row_tris([0, 2]), 8).n_nodes == 1
    t = bvh_oracle.build_morton(*_row_tris(range(8)), 4)
    assert t.n_nodes == 3 and t.left[0] > 0 and t.right[0] > 0
    v0, _, _ = _row_tris([7, 1, 6, 0, 5, 2, 4, 3])
    t = bvh_oracle.build_morton(*_row_tris([7, 1, 6, 0, 5, 2, 4, 3]), 1)
    assert v0[t.order][:, 0].tolist() == [0, 1, 2, 3, 4, 5, 6, 7]


def test_bvh_median_partition_invariant():
    # tests/bvh_builder_tests.py:47-67
    t = bvh_oracle.build_median(*_row_tris([7, 1, 6, 0, 5, 2, 4, 3]), 1)
    for i in range(t.n_nodes):
        if t.left[i] < 0:
            continue
        axis = bvh_oracle._longest_axis(t.aabb_min[i], t.aabb_max[i])
        assert t.aabb_max[t.left[i]][axis] <= t.aabb_max[t.right[i]][axis]


def _interleave(x, y, z):
    code = 0
    for b in range(10):
        code |= ((int(x) >> b) & 1) << (3 * b + 2)
        code |= ((int(y) >> b) & 1) << (3 * b + 1)
        code |= ((int(z) >> b) & 1) << (3 * b)
    return code


@pytest.mark.parametrize("n,leaf", [(1, 8), (37, 8), (1000, 4), (5000, 1)])
def test_morton_tree_invariants(n, leaf):
    rng = np.random.default_rng(n)
    v0 = rng.uniform(-3, 3, size=(n, 3)).astype(np.float32)
    v1 = (v0 + rng.normal(size=(n, 3))).astype(np.float32)
    v2 = (v0 + rng.normal(size=(n, 3))).astype(np.float32)
    t = bvh_oracle.build_morton(v0, v1, v2, leaf)
    lo, hi, cen = bvh_oracle.padded_bounds(v0, v1, v2)
    # codes by an independent bit-by-bit interleave, stable order
    q = np.clip((cen - cen.min(0)) / np.where(np.ptp(cen, 0) <= 1e-12, 1, np.ptp(cen, 0)).astype(np.float32), 0, 1)
    q = (q * np.float32(1023)).astype(np.uint32)
    codes = np.array([_interleave(*row) for row in q], dtype=np.int64)
    assert sorted(range(n), key=lambda i: (codes[i], i)) == t.order.tolist()
    # closed-form shape and preorder links
    assert (t.n_nodes, t.depth) == bvh_oracle.closed_form_shape(n, leaf)
    interior = t.left >= 0
    assert (t.left[interior] == np.nonzero(interior)[0] + 1).all()
    # leaves tile [0, n) in order; each box is the exact union of its triangles
    leaves = np.nonzero(~interior)[0]
    starts = t.tri_start[leaves]
    assert starts[0] == 0 and (np.diff(starts) == t.tri_count[leaves][:-1]).all()
    assert t.tri_count[leaves].sum() == n and t.tri_count[leaves].max() <= leaf

    def span(i):
        if t.left[i] < 0:
            return t.tri_start[i], t.tri_start[i] + t.tri_count[i]
        return span(t.left[i])[0], span(t.right[i])[1]

    for i in rng.choice(t.n_nodes, size=min(200, t.n_nodes), replace=False):
        s, e = span(int(i))
        idx = t.order[s:e]
        np.testing.assert_array_equal(t.aabb_min[i], lo[idx].min(0))
        np.testing.assert_array_equal(t.aabb_max[i], hi[idx].max(0))


def test_object_median_rule_does_not_forward_leaf_size():
    # core/bvh.py:12-13: children use the default leaf size 4
    v0, v1, v2 = _row_tris(range(40))
    t = bvh_oracle.build_object_median(v0, v1, v2, max_leaf=16)
    assert t.left[0] >= 0
    assert t.tri_count[t.tri_count > 0].max() <= 4


def test_object_median_oracle_matches_reference_bvhnode():
    """The oracle's core/bvh.py restatement against the reference's own
    BVHNode run on seeded meshes (tests/golden/object_median.npz, written by
    tests/golden/make_object_median_golden.py): same preorder tree, leaf
    order and fp64 boxes (stored rounded to fp32), including centroid ties
    on a coarse grid, longest-axis ties on a flat mesh and -0.0 / +0.0
    centroids."""
    from pathlib import Path

    g = np.load(Path(__file__).resolve().parent / "golden" / "object_median.npz")
    for k, (n, leaf, _seed) in enumerate(g["cases"]):
        t = bvh_oracle.build_object_median(g[f"case{k}_v0"], g[f"case{k}_v1"], g[f"case{k}_v2"], max_leaf=int(leaf))
        assert t.n_nodes == len(g[f"case{k}_left"]), k
        np.testing.assert_array_equal(t.order, g[f"case{k}_order"])
        for f in ("left", "right", "tri_start", "tri_count"):
            np.testing.assert_array_equal(getattr(t, f), g[f"case{k}_{f}"], err_msg=f"{k} {f}")
        for f in ("aabb_min", "aabb_max"):
            np.testing.assert_array_equal(getattr(t, f), g[f"case{k}_{f}"].astype(np.float32), err_msg=f"{k} {f}")


def test_material_palette_first_appearance():
    rows = [(0, (1, 1, 1), 0, 1, 0, -1), (0, (0.5, 0.5, 0.5), 0, 1, 0, -1), (0, (1, 1, 1), 0, 1, 0, -1)]
    pal, idx = bvh_oracle.material_palette(rows)
    assert len(pal) == 2 and idx.tolist() == [0, 1, 0]


@pytest.mark.parametrize("flip_v", [True, False])
def test_oracle_texture_lookup_closed_form(flip_v):
    """trace.py:96-136 in the oracle: on a floor with uv = world (x, z) and a
    texture linear in its texel coordinates, the first-hit albedo is the
    linear map at the hit's wrapped (and flipped) uv, for every camera ray."""
    from roxy_dcc_amd.extract import flatten_scene

    W, H = 16, 8
    world, tex, _ = helpers.textured_floor_world(W, H, flip_v=flip_v)
    flat = flatten_scene(world)
    assert len(flat.textures) == 1
    osc = helpers.flat_to_oracle(flat)
    cam = world.active_camera
    ocam = helpers.oracle_camera(cam, 48, 48)
    acc = orc.Accum(48 * 48)
    orc.render(osc, ocam, orc.params(1, 0, 0, 1, 10.0, (0.5, 0.5, 0.5), 3), 0, 1, acc)
    rays = orc.camera_rays(ocam, 0)
    o = np.asarray(cam.position, np.float64)
    t = acc.depth.astype(np.float64)
    hit = t > 0
    assert hit.mean() > 0.3
    p = o[None, :] + rays.astype(np.float64) * t[:, None]
    want = helpers.textured_albedo_closed_form(p[hit], W, H, flip_v=flip_v)
    got = acc.albedo[hit].astype(np.float64)
    # away from the wrap seams (where uv, computed in fp32, may land on either side)
    u, v = p[hit, 0] % 1.0, p[hit, 2] % 1.0
    inner = (np.minimum(u, 1 - u) > 1e-3) & (np.minimum(v, 1 - v) > 1e-3)
    assert inner.mean() > 0.9
    assert np.max(np.abs(got[inner] - want[inner])) < 2e-4


def test_fp32_bias_on_large_scene_is_the_algorithms_own():
    """The fp64 fixture of the 2M-triangle scene is reproduced without bias by
    an independent fp64 render, while the fp32 oracle (the arithmetic the
    GPU matches bit for bit) sits within 0.2% of the mean radiance of it."""
    from pathlib import Path

    from roxy_dcc_amd import scenes
    from roxy_dcc_amd.extract import flatten_scene

    ref = np.load(Path(__file__).resolve().parent / "golden" / "rmse_large.npz")
    sel = np.sort(np.random.default_rng(1).choice(len(ref["pixels"]), 2048, replace=False))
    pix, want = ref["pixels"][sel], ref["color"][sel]
    world, W, H, st = scenes.large_scene()
    osc = helpers.flat_to_oracle(flatten_scene(world))
    ocam = helpers.oracle_camera(world.active_camera, W, H)
    prm = orc.params(st["max_depth"], 1, 0, st["direct_light_max_depth"], 10.0, world._background_color, 11)
    out = {}
    for prec in ("f64", "f32"):
        acc = orc.Accum(len(pix), prec)
        orc.render(osc, ocam, prm, 0, 256, acc, pixels=pix)
        err = np.asarray(acc.color, np.float64) - want
        out[prec] = (err.mean(), err.std() / np.sqrt(err.size))
    assert abs(out["f64"][0]) < 5 * out["f64"][1]
    assert abs(out["f32"][0]) < 5 * out["f32"][1] + 2e-3 * float(np.mean(want))
