"""Generate the fp64 reference renders used for RMSE statements.

For each configuration, a seeded subset of pixels is rendered by the CPU
oracle's fp64 instance (the reference estimator of
rendering/taichi/trace.py in double precision) at 4096 samples per pixel,
and the per-pixel means of the samples and of their squares are stored
with the pixel indices:

    tests/golden/rmse_<config>.npz  {pixels, color (n,3) f64, color_sq (n,3) f64, spp, seed, ...}

Run from the repository root:  python tests/golden/make_rmse_refs.py [config ...]
The GPU tests compare renders of the same pixels against these means.
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as orc  # noqa: E402
from roxy_dcc_amd import scenes  # noqa: E402
from roxy_dcc_amd.extract import flatten_scene  # noqa: E402
from roxy_dcc_amd.renderer import TaichiRenderer  # noqa: E402
from tests import helpers  # noqa: E402

N_PIXELS = {"cornell": 4096, "mesh": 4096, "large": 65536}
SPP = 4096
PIXEL_SEED = 1234
RENDER_SEED = 0x5EED


def make(name, n_pixels=None, spp=SPP):
    n_pixels = n_pixels or N_PIXELS[name]
    world, W, H, st = scenes.CONFIGS[name]()
    flat = flatten_scene(world)
    osc = helpers.flat_to_oracle(flat)
    cam = world.active_camera
    ocam = helpers.oracle_camera(cam, W, H)
    mode = TaichiRenderer._direct_light_mode_to_id(st["direct_light_mode"])
    prm = orc.params(st["max_depth"], int(world._use_sky), mode, st["direct_light_max_depth"], 10.0,
                     world._background_color, RENDER_SEED)
    pix = np.sort(np.random.default_rng(PIXEL_SEED).choice(W * H, size=n_pixels, replace=False)).astype(np.int32)
    acc = orc.Accum(n_pixels, "f64", moments=True)
    t0 = time.time()
    orc.render(osc, ocam, prm, 0, spp, acc, pixels=pix)
    dt = time.time() - t0
    out = ROOT / "tests" / "golden" / f"rmse_{name}.npz"
    np.savez_compressed(out, pixels=pix, color=acc.color, color_sq=acc.sq / spp, spp=spp, render_seed=RENDER_SEED,
                        width=W, height=H, max_depth=st["max_depth"])
    print(f"{name}: {n_pixels} px x {spp} spp in {dt:.0f} s -> {out.name}")


if __name__ == "__main__":
    for name in (sys.argv[1:] or ["cornell", "mesh", "large"]):
        make(name)
