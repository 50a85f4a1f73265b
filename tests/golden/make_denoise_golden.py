"""Golden vectors of the reference's edge_aware_denoise, from its own code.

rendering/denoise.py imports numpy only; it is loaded here by file path
(rendering/__init__.py would pull in pygame/moderngl) and run on seeded
noisy images with every combination of guide buffers (none, normal,
albedo, depth as HxW and as HxWx1, pairs, all three), radii 1-3, several
sigma/amount settings, and the identity cases (radius 0, amount 0).

Inputs are shared per image size; each case stores its settings and the
output:

    tests/golden/denoise.npz
      img<k>_color, img<k>_normal, img<k>_albedo, img<k>_depth   inputs
      case<c>_out                                                output
      cases: (image, radius, guides bitmask, depth as HxWx1) per case
      case_sigmas: (sigma_color, amount, sigma_normal, sigma_albedo, sigma_depth)

Run from the repository root, in a container that holds the reference:
    python tests/golden/make_denoise_golden.py [/path/to/reference]
The tests read only the .npz.
"""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np

sys.dont_write_bytecode = True
ROOT = Path(__file__).resolve().parents[2]
REFERENCE = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

GUIDE_NORMAL, GUIDE_ALBEDO, GUIDE_DEPTH = 1, 2, 4
SIZES = [(13, 17), (24, 32)]


def _load_denoise():
    spec = importlib.util.spec_from_file_location("ref_denoise", REFERENCE / "rendering/denoise.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def image(k, h, w):
    """A noisy piecewise-smooth image with its guides: two regions split by a
    diagonal edge (different normal, albedo and depth), colour noise."""
    rng = np.random.default_rng(50 + k)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float32)
    side = (xs * 0.7 + ys > 0.6 * w).astype(np.float32)[..., None]
    base = side * np.float32([0.8, 0.5, 0.2]) + (1 - side) * np.float32([0.1, 0.3, 0.6])
    color = (base + rng.normal(scale=0.12, size=(h, w, 3))).astype(np.float32)
    n_a = np.float32([0, 0, 1])
    n_b = np.float32([0.6, 0, 0.8])
    normal = (side * n_b + (1 - side) * n_a + rng.normal(scale=0.02, size=(h, w, 3))).astype(np.float32)
    albedo = (side * np.float32([0.9, 0.9, 0.9]) + (1 - side) * np.float32([0.3, 0.5, 0.4])).astype(np.float32)
    depth = (side[..., 0] * 3.0 + 1.0 + xs * 0.05 + rng.normal(scale=0.01, size=(h, w))).astype(np.float32)
    return color, normal, albedo, depth


def cases():
    """(image, radius, guides, depth_3d, sigma_color, amount, sigma_n, sigma_a, sigma_d)"""
    out = []
    for img in range(len(SIZES)):
        for radius in (1, 2, 3):
            for guides in range(8):
                out.append((img, radius, guides, 0, 0.08, 0.8, 0.35, 0.20, 1.0))
        out.append((img, 1, GUIDE_DEPTH, 1, 0.08, 0.8, 0.35, 0.20, 1.0))
        out.append((img, 2, 7, 1, 0.3, 1.0, 0.1, 0.5, 0.25))
        out.append((img, 1, 0, 0, 1e-5, 0.5, 0.35, 0.20, 1.0))
        out.append((img, 1, 7, 0, 0.08, 1.7, 0.35, 0.20, 1.0))  # amount clipped to 1
        out.append((img, 0, 7, 0, 0.08, 0.8, 0.35, 0.20, 1.0))  # radius 0: identity
        out.append((img, 2, 7, 0, 0.08, 0.0, 0.35, 0.20, 1.0))  # amount 0: identity
    return out


def main():
    mod = _load_denoise()
    out = {}
    for k, (h, w) in enumerate(SIZES):
        for name, arr in zip(("color", "normal", "albedo", "depth"), image(k, h, w)):
            out[f"img{k}_{name}"] = arr
    cs = cases()
    out["cases"] = np.asarray([c[:4] for c in cs], np.int32)
    out["case_sigmas"] = np.asarray([c[4:] for c in cs], np.float64)
    for c, (img, radius, guides, d3, sc, amount, sn, sa, sd) in enumerate(cs):
        color = out[f"img{img}_color"]
        depth = out[f"img{img}_depth"]
        kw = {}
        if guides & GUIDE_NORMAL:
            kw["normal"] = out[f"img{img}_normal"]
        if guides & GUIDE_ALBEDO:
            kw["albedo"] = out[f"img{img}_albedo"]
        if guides & GUIDE_DEPTH:
            kw["depth"] = depth[..., None] if d3 else depth
        res = mod.edge_aware_denoise(color, radius, sc, amount, sigma_normal=sn, sigma_albedo=sa, sigma_depth=sd,
                                     **kw)
        out[f"case{c}_out"] = np.asarray(res, np.float32)
    out["display_in"] = np.float32([-1.0, 0.0, 0.25, 0.5, 1.0, 4.0, np.float32(1e-8)])
    out["display_out"] = np.asarray(mod.linear_to_display(out["display_in"]), np.float32)
    np.savez_compressed(ROOT / "tests" / "golden" / "denoise.npz", **out)
    print(f"{len(cs)} denoise cases over {len(SIZES)} images")


if __name__ == "__main__":
    main()
