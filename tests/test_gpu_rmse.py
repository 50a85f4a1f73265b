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


@pytest.mark.parametrize("name,spp", [("cornell", 64), ("mesh", 64), ("large", 16)])
def test_pixels_within_three_sigma_of_fp64_oracle(name, spp):
    """Per pixel and channel, the render against the fp64 oracle's 4096-spp
    mean in units of the combined standard error sqrt(v / spp + v / 4096),
    v the per-sample variance the oracle measured at that pixel (its second
    moment minus the squared mean).  Pixels where all 4096 reference samples
    were equal (v = 0: background, unlit or fully specular-to-background)
    must match the reference value.  Gaussian errors would put 99.7% within
    3 sigma.  Pixels whose variance comes from rare bright samples (implicit
    light hits, clamped at 10) are not Gaussian at 64 samples: drawing one
    such sample moves the mean by more than 3 sigma.  Measured: cornell
    98.4%, mesh 98.9% (64 spp), large 99.4% (16 spp); the render equals the
    fp32 oracle's, so these figures are deterministic.  The bound is 98%."""
    ref = np.load(GOLDEN / f"rmse_{name}.npz")
    pix, want = ref["pixels"], ref["color"]
    var = np.maximum(ref["color_sq"] - want ** 2, 0.0)
    got = _gpu_render(name, spp, seed=11)[pix].astype(np.float64)
    sig = np.sqrt(var / spp + var / float(ref["spp"]))
    spread = sig > 1e-9 * (1.0 + np.abs(want))
    z = np.abs(got - want)[spread] / sig[spread]
    frac = float((z <= 3).mean())
    flat_ok = float((np.abs(got - want)[~spread] <= 1e-5 * (1.0 + np.abs(want[~spread]))).mean()) if (~spread).any() else 1.0
    print(f"{name}: {frac * 100:.2f}% of {z.size} pixel channels within 3 sigma at {spp} spp "
          f"(99th percentile {np.quantile(z, 0.99):.2f} sigma); {int((~spread).sum())} without variance, "
          f"{flat_ok * 100:.2f}% of them equal", flush=True)
    assert frac >= 0.98
    assert flat_ok >= 0.99
