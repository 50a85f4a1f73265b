"""rx_render_noalloc: the render with its workspace in the library's own
device array.  Two calls in a row (the second with a smaller frame and
pass) give bit-identical accumulators, display and counters to rx_render
with a caller workspace; a pass larger than the array is rejected before
any GPU work."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests import helpers

pytestmark = pytest.mark.gpu


def _setup(name, W, H):
    from roxy_dcc_amd.extract import extract_scene

    world, _, _, st = helpers.load(name)
    return world, st, extract_scene(world).scene


def _params(R, world, st, spp_pass, seed):
    mode = R.TaichiRenderer._direct_light_mode_to_id(st["direct_light_mode"])
    return R.params_struct(st["max_depth"], int(world._use_sky), mode, st["direct_light_max_depth"], 10.0,
                           world._background_color, seed, spp_pass)


def _both(name, W, H, samples, spp_pass, seed):
    import torch

    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd import renderer as R

    world, st, sc = _setup(name, W, H)
    cam = R.camera_struct(world.active_camera, W, H)
    prm = _params(R, world, st, spp_pass, seed)
    integ = R.Integrator(sc, W, H, spp_pass)
    ref = R.DeviceFrame(W, H)
    got = R.DeviceFrame(W, H)
    ctr = torch.zeros(8, dtype=torch.int64, device="cuda")
    done = 0
    while done < samples:
        n = min(spp_pass, samples - done)
        integ.render(ref, cam, prm, done, n)
        nat.check(nat.lib().rx_render_noalloc(C.byref(sc.struct()), C.byref(cam), C.byref(prm), done, n,
                                              C.byref(got.struct()), ctr.data_ptr(), nat.stream_ptr()),
                  "render noalloc")
        done += n
    torch.cuda.synchronize()
    for k in ("color", "normal", "albedo", "depth", "display"):
        np.testing.assert_array_equal(getattr(got, k).cpu().numpy(), getattr(ref, k).cpu().numpy(), err_msg=k)
    assert ctr.cpu().tolist() == integ.counters.cpu().tolist()
    return int(ctr[0].item())


def test_noalloc_render_equals_workspace_render_twice():
    print("no-alloc call 1: mesh 512x512, 8 samples in passes of 4", flush=True)
    rays1 = _both("mesh", 512, 512, 8, 4, seed=3)
    print(f"  rays {rays1}", flush=True)
    print("no-alloc call 2: cornell 64x48, 3 samples in passes of 2", flush=True)
    rays2 = _both("cornell", 64, 48, 3, 2, seed=4)
    print(f"  rays {rays2}", flush=True)
    assert rays1 > rays2 > 0


def test_noalloc_rejects_passes_beyond_the_array():
    import torch

    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd import renderer as R

    world, st, sc = _setup("large", 1920, 1080)
    cam = R.camera_struct(world.active_camera, 1920, 1080)
    frame = R.DeviceFrame(1920, 1080)
    ctr = torch.zeros(8, dtype=torch.int64, device="cuda")
    need = C.c_size_t()
    nat.check(nat.lib().rx_render_workspace_size(1920, 1080, 2, sc.builder.depth, C.byref(need)))
    assert need.value > nat.RX_NOALLOC_SCRATCH_BYTES
    st_code = nat.lib().rx_render_noalloc(C.byref(sc.struct()), C.byref(cam), C.byref(_params(R, world, st, 2, 0)),
                                          0, 2, C.byref(frame.struct()), ctr.data_ptr(), nat.stream_ptr())
    assert st_code == nat.RX_ERR_CAPACITY
    assert "rx_render_noalloc has" in nat.last_error()
    prof = _params(R, world, st, 1, 0)
    prof.flags |= nat.RX_FLAG_PROFILE
    assert nat.lib().rx_render_noalloc(C.byref(sc.struct()), C.byref(cam), C.byref(prof), 0, 1,
                                       C.byref(frame.struct()), ctr.data_ptr(), nat.stream_ptr()) == nat.RX_ERR_INVALID
    torch.cuda.synchronize()
    assert ctr.cpu().tolist() == [0] * 8 and float(frame.color.abs().sum()) == 0.0
