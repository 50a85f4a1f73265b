"""Error against the fp64 oracle at 4096 spp (tests/golden/rmse_*.npz).

The GPU render (independent seed) must be unbiased against the fp64
reference means on the fixture's pixels up to fp32's own systematic error
(0.2% of the mean radiance), and its RMSE must fall as 1/sqrt(spp).  The printed RMSE values are the per-configuration figures."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from tests import helpers

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _gpu_render(name, spp, seed):
    import torch

    from roxy_dcc_amd import renderer as R
    from roxy_dcc_amd.extract import extract_scene

    world, W, H, st = helpers.load(name)
    ex = extract_scene(world)
    per = R.default_samples_per_pass(W, H, spp)
    integ = R.Integrator(ex.scene, W, H, per)
    frame = R.DeviceFrame(W, H)
    cam = R.camera_struct(world.active_camera, W, H)
    mode = R.TaichiRenderer._direct_light_mode_to_id(st["direct_light_mode"])
    prm = R.params_struct(st["max_depth"], int(world._use_sky), mode, st["direct_light_max_depth"], 10.0,
                          world._background_color, seed, per)
    done = 0
    while done < spp:
        integ.render(frame, cam, prm, done, min(per, spp - done))
        done += per
    torch.cuda.synchronize()
    return frame.color.cpu().numpy().reshape(-1, 3)


@pytest.mark.parametrize("name", ["cornell", "mesh", "large"])
def test_rmse_against_fp64_oracle(name):
    ref = np.load(GOLDEN / f"rmse_{name}.npz")
    pix, want = ref["pixels"], ref["color"]
    rm = {}
    for spp in (16, 64):
        got = _gpu_render(name, spp, seed=11)[pix].astype(np.float64)
        err = got - want
        rm[spp] = float(np.sqrt(np.mean(err ** 2)))
        bias = float(err.mean())
        se = float(err.std() / np.sqrt(err.size))
        print(f"{name}: {spp} spp RMSE {rm[spp]:.5f} (vs fp64 oracle at {int(ref['spp'])} spp), "
              f"mean error {bias:.2e} +- {se:.1e}", flush=True)
        # fp32 arithmetic of the reference algorithm has a small systematic
        # error of its own (on the 2M-triangle scene the fp32 oracle is
        # 0.08% darker than the fp64 one: test_oracle_pins checks this), so
        # the bound allows 0.2% of the mean radiance beyond the noise
        assert abs(bias) < 5 * se + 1e-4 + 2e-3 * float(np.mean(np.abs(want)))
    assert 1.5 < rm[16] / rm[64] < 2.6
