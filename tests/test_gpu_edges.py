"""Edge cases of the device path against the oracle: an empty world, depth
0, ragged image sizes, a world of analytic primitives only, and hit queries
of zero and one ray.

The ray-count rules are those of tests/ray_tracer_tests.py:107-122 (an empty
scene casts W*H*spp rays; depth 0 casts none)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from tests import helpers

pytestmark = pytest.mark.gpu

SETTINGS = dict(samples=3, max_depth=4, direct_light_mode="one", direct_light_max_depth=1)


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _assert_same(res, label):
    for k in ("color", "normal", "albedo", "depth"):
        np.testing.assert_array_equal(res[f"gpu_{k}"], res[f"cpu_{k}"], err_msg=f"{label}: {k}")
    assert res["gpu_rays"] == res["cpu_rays"], label


def _empty_world(use_sky, bg=(0.2, 0.3, 0.4)):
    from roxy_dcc_amd.world import Camera, World

    w = World(background_color=bg, use_sky=use_sky)
    w.add_camera(Camera(position=(0.0, 0.0, 0.0), forward=(0.3, 0.2, 1.0), fov=50, width=24, height=16))
    return w


@pytest.mark.parametrize("use_sky", [True, False])
def test_empty_world_casts_one_ray_per_sample(use_sky):
    _torch()
    W, H, spp = 24, 16, 3
    res = helpers.render_pair(None, samples=spp, world_tuple=(_empty_world(use_sky), W, H, dict(SETTINGS)))
    _assert_same(res, f"empty sky={use_sky}")
    assert res["gpu_rays"] == W * H * spp
    if not use_sky:
        np.testing.assert_array_equal(res["gpu_color"], np.tile(np.float32([0.2, 0.3, 0.4]), (W * H, 1)))


def test_depth_zero_casts_no_rays():
    _torch()
    res = helpers.render_pair("cornell", width=20, height=20, samples=2, max_depth=0)
    _assert_same(res, "depth 0")
    assert res["gpu_rays"] == 0
    assert not res["gpu_color"].any()


@pytest.mark.parametrize("W,H", [(1, 1), (37, 23), (3, 131), (130, 2)])
def test_ragged_image_sizes(W, H):
    _torch()
    res = helpers.render_pair("cornell", width=W, height=H, samples=3)
    _assert_same(res, f"cornell {W}x{H}")
    assert res["gpu_color"].shape == (W * H, 3)


def test_primitives_only_world():
    """No triangles at all: the traversal sees an empty BVH and every hit
    comes from the primitive scan (intersect.py:259-287)."""
    from roxy_dcc_amd.world import (Camera, Cube, Dielectric, Diffuse, Emissive, Mat4, Metal, Plane, SceneObject,
                                    Shape, Sphere, World)

    _torch()
    w = World(background_color=(0.05, 0.05, 0.05), use_sky=True)
    w.add_object(SceneObject("floor", [Shape(Plane((0.0, 1.0, 0.0), 0.0), Diffuse((0.7, 0.7, 0.7)))], Mat4()))
    w.add_object(SceneObject("ball", [Shape(Sphere(0.5), Metal((0.8, 0.6, 0.3), 0.1))],
                             Mat4.trs(translation=(0.0, 0.5, 3.0))))
    w.add_object(SceneObject("glass", [Shape(Sphere(0.4), Dielectric(1.5))], Mat4.trs(translation=(1.0, 0.4, 2.5))))
    w.add_object(SceneObject("box", [Shape(Cube(0.6), Diffuse((0.2, 0.5, 0.8)))],
                             Mat4.trs(translation=(-1.0, 0.3, 3.0))))
    w.add_object(SceneObject("light", [Shape(Sphere(0.3), Emissive((1.0, 1.0, 1.0), 10.0))],
                             Mat4.trs(translation=(0.0, 2.5, 2.0))))
    w.add_camera(Camera(position=(0.0, 1.0, -1.0), forward=(0.0, -0.2, 1.0), fov=60, width=40, height=30))
    res = helpers.render_pair(None, samples=4, world_tuple=(w, 40, 30, dict(SETTINGS)))
    assert res["extract"].flat.triangles().triangle_count == 0
    _assert_same(res, "primitives only")
    assert res["gpu_counts"]["shadow_rays"] > 0


def test_zero_and_one_ray_queries():
    """n = 0 returns at once and writes nothing; n = 1 matches the oracle."""
    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd.extract import extract_scene

    torch = _torch()
    world, *_ = helpers.load("cornell")
    ex = extract_scene(world)
    osc = helpers.flat_to_oracle(ex.flat)
    lib, scene = nat.lib(), ex.scene.struct()
    for n in (0, 1):
        rays = helpers.camera_rays_np(world.active_camera, 1, 1)[:, :n].copy()
        d_rays = torch.from_numpy(rays).cuda() if n else torch.empty((6, 1), device="cuda")
        out = {k: torch.full((1,), 7, dtype=torch.float32 if k in ("t", "u", "v") else torch.int32, device="cuda")
               for k in ("t", "id", "kind", "u", "v")}
        nat.check(lib.rx_closest_hit(C.byref(scene), d_rays.data_ptr(), n, out["t"].data_ptr(),
                                     out["id"].data_ptr(), out["kind"].data_ptr(), out["u"].data_ptr(),
                                     out["v"].data_ptr(), None, nat.stream_ptr()))
        tmax = torch.full((1,), 100.0, device="cuda")
        occ = torch.full((1,), 7, dtype=torch.int32, device="cuda")
        nat.check(lib.rx_any_hit(C.byref(scene), d_rays.data_ptr(), tmax.data_ptr(), n, occ.data_ptr(), None,
                                 nat.stream_ptr()))
        torch.cuda.synchronize()
        got = {k: v.cpu().numpy() for k, v in out.items()}
        if n == 0:
            assert all((v == 7).all() for v in got.values()) and int(occ.item()) == 7
            continue
        want = orc.closest_hit(osc, rays)
        assert got["kind"][0] == want["kind"][0] and got["id"][0] == want["id"][0]
        assert got["t"].view(np.uint32)[0] == want["t"].view(np.uint32)[0]
        assert int(occ.item()) == int(orc.any_hit(osc, rays, np.float32([100.0]))[0])


def test_negative_ray_count_is_an_error():
    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd.extract import extract_scene

    torch = _torch()
    world, *_ = helpers.load("cornell")
    ex = extract_scene(world)
    buf = torch.zeros(8, device="cuda")
    st = nat.lib().rx_closest_hit(C.byref(ex.scene.struct()), buf.data_ptr(), -1, buf.data_ptr(), buf.data_ptr(),
                                  buf.data_ptr(), buf.data_ptr(), buf.data_ptr(), None, nat.stream_ptr())
    assert st != 0
    with pytest.raises(Exception):
        nat.check(st)


@pytest.mark.parametrize("kind", ["cornell", "root_leaf"])
def test_treelet_on_trees_smaller_than_the_treelet(kind):
    """Passes of 2^19 path slots or more trace with the LDS copy of the top
    of the tree.  A tree with fewer interior nodes than the treelet holds
    (the Cornell box: 15 nodes) goes into it whole; a tree whose root is a
    leaf (two triangles) gives an empty treelet and the traversal starts
    from the root record."""
    from roxy_dcc_amd.world import Camera, Diffuse, Emissive, IndexedMesh, Mat4, SceneObject, Shape, Sphere, World

    _torch()
    W, H = 768, 704    # 540,672 paths at one sample: above 2^19
    if kind == "cornell":
        res = helpers.render_pair("cornell", width=W, height=H, samples=1)
    else:
        w = World(background_color=(0.1, 0.1, 0.1), use_sky=False)
        pos = np.array([[-2, 0, -2], [2, 0, -2], [2, 0, 2], [-2, 0, 2]], np.float64)
        w.add_object(SceneObject("floor", [Shape(IndexedMesh(pos, np.int32([[0, 2, 1], [0, 3, 2]])),
                                                 Diffuse((0.6, 0.5, 0.4)))], Mat4()))
        w.add_object(SceneObject("light", [Shape(Sphere(0.3), Emissive((1.0, 1.0, 1.0), 10.0))],
                                 Mat4.trs(translation=(0.0, 1.5, 0.0))))
        w.add_camera(Camera(position=(0.0, 2.0, -4.0), forward=(0.0, -0.4, 1.0), fov=50, width=W, height=H))
        res = helpers.render_pair(None, samples=1, world_tuple=(w, W, H, dict(SETTINGS)))
        assert res["extract"].flat.triangles().triangle_count == 2
    _assert_same(res, f"treelet {kind}")
