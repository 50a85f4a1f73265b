"""The camera-ray traversal on the scene's eye-relative copies (node boxes
and triangle vertices minus the camera position, with the test terms built
from them alone) renders exactly what the traversal that recomputes them
per ray renders: every accumulator bit, the ray count and every traversal
counter (node pops, triangle and primitive tests), on scenes with triangle
meshes, spheres, planes and cubes, in both next-event modes."""
from __future__ import annotations

import numpy as np
import pytest

from roxy_dcc_amd import _native as nat
from tests import helpers

pytestmark = pytest.mark.gpu


def _render(world, W, H, settings, samples, spp, eye, seed=11, mode=None):
    import torch

    from roxy_dcc_amd import renderer as R
    from roxy_dcc_amd.extract import extract_scene

    ex = extract_scene(world)
    sc = ex.scene.struct()
    records = (sc.n_nodes + sc.n_tris if sc.n_tris > 0 else 0) + sc.n_prims
    # rx_render takes the copies when a call's camera rays are at least four per record
    assert W * H * min(spp, samples) >= 4 * records, (W, H, spp, records)
    integ = R.Integrator(ex.scene, W, H, spp, eye_copies=eye)
    frame = R.DeviceFrame(W, H)
    cam = R.camera_struct(world.active_camera, W, H)
    mode_id = R.TaichiRenderer._direct_light_mode_to_id(mode or settings["direct_light_mode"])
    prm = R.params_struct(settings["max_depth"], int(world._use_sky), mode_id, settings["direct_light_max_depth"],
                          10.0, world._background_color, seed, spp, flags=nat.RX_FLAG_COUNT_TRAFFIC)
    done = 0
    while done < samples:
        n = min(spp, samples - done)
        integ.render(frame, cam, prm, done, n)
        done += n
    torch.cuda.synchronize()
    out = {k: getattr(frame, k).cpu().numpy() for k in ("color", "normal", "albedo", "depth")}
    return out, integ.counts(), integ.workspace.numel()


def _same(world, W, H, settings, samples, spp, **kw):
    a, ca, wa = _render(world, W, H, settings, samples, spp, True, **kw)
    b, cb, wb = _render(world, W, H, settings, samples, spp, False, **kw)
    assert wa > wb, "the scene-sized workspace holds the eye-relative copies"
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert ca == cb, (ca, cb)
    assert ca["node_pops"] + ca["prim_tests"] > 0


@pytest.mark.parametrize("name", ["cornell", "mesh"])
def test_eye_copies_exact_on_configs(name):
    world, W, H, st = helpers.load(name)
    W, H = (96, 72) if name == "cornell" else (512, 512)
    print(f"eye copies: {name} at {W}x{H}", flush=True)
    _same(world, W, H, st, 4, 2 if name == "cornell" else 4)


def test_eye_copies_exact_on_reduced_large_scene():
    from roxy_dcc_amd import scenes

    world, W, H, st = scenes.large_scene(seed=30, n_meshes=40, seg_u=20, seg_v=11, n_spheres=16)
    print("eye copies: reduced large scene at 384x216", flush=True)
    _same(world, 384, 216, st, 3, 3)


@pytest.mark.parametrize("seed", [3, 8, 21])
def test_eye_copies_exact_on_random_worlds(seed):
    from tests.test_gpu_random_scenes import random_world

    for mode in ("one", "all"):
        world, W, H, st = random_world(seed, mode)
        print(f"eye copies: random world {seed} mode {mode} at 160x112", flush=True)
        _same(world, 160, 112, st, 3, 3, mode=mode)
