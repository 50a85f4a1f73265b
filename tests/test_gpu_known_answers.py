"""The reference's geometry known answers, traced on the device.

The same cases the reference's CPU tests assert (tests/mesh_tests.py:13-43,
105-112: centre hit t = 1, misses, barycentrics in range, closest of two
t = 2; tests/shapes_tests.py:53-105: sphere t = 4 from outside and t = 1
from inside; tests/aabb_tests.py:11-28: slab t = 1) go through
rx_closest_hit / rx_any_hit on scenes made with the world API, and every
result also equals the oracle's bit for bit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from tests import helpers

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _world(tris=(), prims=()):
    from roxy_dcc_amd.world import Camera, Diffuse, IndexedMesh, Mat4, SceneObject, Shape, World

    w = World(background_color=(0.0, 0.0, 0.0), use_sky=False)
    if len(tris):
        t = np.asarray(tris, np.float64).reshape(-1, 3, 3)
        mesh = IndexedMesh(t.reshape(-1, 3), np.arange(3 * len(t), dtype=np.int32).reshape(-1, 3))
        w.add_object(SceneObject("mesh", [Shape(mesh, Diffuse((0.5, 0.5, 0.5)))], Mat4()))
    for k, (shape, at) in enumerate(prims):
        w.add_object(SceneObject(f"prim{k}", [Shape(shape, Diffuse((0.5, 0.5, 0.5)))], Mat4.trs(translation=at)))
    w.add_camera(Camera())
    return w


def _rays(o, d):
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    return np.ascontiguousarray(np.concatenate([o.T, d.T]))


def _trace(world, rays, tmax=None):
    """(device closest hit, oracle closest hit[, device any hit, oracle any hit])."""
    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd.extract import extract_scene

    torch = _torch()
    ex = extract_scene(world)
    osc = helpers.flat_to_oracle(ex.flat)
    n = rays.shape[1]
    d_rays = torch.from_numpy(rays).cuda()
    out = {k: torch.empty(n, dtype=torch.float32 if k in ("t", "u", "v") else torch.int32, device="cuda")
           for k in ("t", "id", "kind", "u", "v")}
    scene = ex.scene.struct()
    nat.check(nat.lib().rx_closest_hit(C.byref(scene), d_rays.data_ptr(), n, out["t"].data_ptr(),
                                       out["id"].data_ptr(), out["kind"].data_ptr(), out["u"].data_ptr(),
                                       out["v"].data_ptr(), None, nat.stream_ptr()))
    res = [None, orc.closest_hit(osc, rays)]
    if tmax is not None:
        d_t = torch.from_numpy(np.asarray(tmax, np.float32)).cuda()
        occ = torch.empty(n, dtype=torch.int32, device="cuda")
        nat.check(nat.lib().rx_any_hit(C.byref(scene), d_rays.data_ptr(), d_t.data_ptr(), n, occ.data_ptr(), None,
                                       nat.stream_ptr()))
        res += [occ, orc.any_hit(osc, rays, np.asarray(tmax, np.float32))]
    torch.cuda.synchronize()
    res[0] = {k: v.cpu().numpy() for k, v in out.items()}
    if tmax is not None:
        res[2] = res[2].cpu().numpy()
    g, o = res[0], res[1]
    np.testing.assert_array_equal(g["kind"], o["kind"])
    np.testing.assert_array_equal(g["id"], o["id"])
    np.testing.assert_array_equal(g["t"].view(np.uint32), o["t"].view(np.uint32))
    return res


UNIT_TRI = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]


def test_triangle_centre_hit_misses_and_barycentrics():
    rays = _rays([[0.25, 0.25, -1], [2, 2, -1], [0.25, 0.25, 0], [0.25, 0.25, 2], [0.3, 0.2, -1]],
                 [[0, 0, 1], [0, 0, 1], [1, 0, 0], [0, 0, 1], [0, 0, 1]])
    g, _, occ, occ_o = _trace(_world([UNIT_TRI]), rays, tmax=[5.0, 5.0, 5.0, 5.0, 0.5])
    assert g["kind"].tolist() == [2, 0, 0, 0, 2]
    assert g["t"][0] == 1.0
    assert (g["t"][1:4] >= 1e9).all() or (g["t"][1:4] < 0).all()
    u, v = g["u"][4], g["v"][4]
    assert u >= 0 and v >= 0 and u + v <= 1
    # occluded only where a hit lies inside (0.001, tmax)
    assert occ.tolist() == [1, 0, 0, 0, 0]
    np.testing.assert_array_equal(occ, occ_o)


def test_closest_of_two_triangles():
    g, *_ = _trace(_world([[[0, 0, 1], [1, 0, 1], [0, 1, 1]], [[0, 0, 3], [1, 0, 3], [0, 1, 3]]]),
                   _rays([[0.25, 0.25, -1]], [[0, 0, 1]]))
    assert g["kind"][0] == 2 and abs(g["t"][0] - 2.0) < 1e-6


def test_sphere_outside_inside_and_misses():
    from roxy_dcc_amd.world import Sphere

    rays = _rays([[0, 0, -5], [0, 0, 0], [0, 5, -5], [0, 0, 5]], [[0, 0, 1]] * 4)
    g, *_ = _trace(_world(prims=[(Sphere(1.0), (0.0, 0.0, 0.0))]), rays)
    assert g["kind"].tolist() == [1, 1, 0, 0]
    assert abs(g["t"][0] - 4.0) < 1e-6 and abs(g["t"][1] - 1.0) < 1e-6


def test_plane_and_cube_slabs():
    from roxy_dcc_amd.world import Cube, Plane

    rays = _rays([[0, 1, 0], [3, 5, 5], [0, 1, 0]], [[0, -1, 0], [1, 0, 0], [1, 0, 0]])
    g, *_ = _trace(_world(prims=[(Plane((0.0, 1.0, 0.0), 0.0), (0.0, 0.0, 0.0)), (Cube(1.0), (5.0, 5.0, 5.0))]),
                   rays)
    assert g["kind"].tolist() == [1, 1, 0]
    assert g["t"][0] == 1.0 and g["t"][1] == 1.5


def test_primitive_wins_a_tie_with_the_bvh():
    """A sphere and a triangle at the same t: the primitive is kept
    (intersect.py:292 takes the BVH hit only if strictly closer)."""
    from roxy_dcc_amd.world import Sphere

    g, *_ = _trace(_world([[[-1, -1, 1], [1, -1, 1], [0, 1, 1]]], prims=[(Sphere(1.0), (0.0, 0.0, 2.0))]),
                   _rays([[0, 0, 0]], [[0, 0, 1]]))
    assert g["kind"][0] == 1 and g["t"][0] == 1.0
