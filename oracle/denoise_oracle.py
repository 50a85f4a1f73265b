"""numpy restatement of the reference edge-aware denoise.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()``
and bench.py's CPU-baseline leg, never by the product package.

Follows rendering/denoise.py:4-108: a (2r+1)^2 window over a reflect-padded
image; weight = 1/(1+dx^2+dy^2) * exp(-|dc|^2 / 2 sc^2) * prod over the
given guides of exp(-|dg|^2 / 2 sg^2); output src*(1-a) + (sum w q / max(sum
w, 1e-8))*a, all in float32.  ``dilation`` spaces the taps for a-trous
passes (dilation 1 is the reference filter exactly).
"""
from __future__ import annotations

import numpy as np


def _guide(name, guide, shape, channels):
    if guide is None:
        return None
    arr = np.asarray(guide, dtype=np.float32)
    if arr.shape[:2] != shape:
        raise ValueError(f"{name} guide must match pixel height and width")
    if arr.ndim == 2:
        arr = arr[:, :, None]
    elif arr.ndim != 3:
        raise ValueError(f"{name} guide must be a 2D or 3D array")
    if arr.shape[2] != channels:
        raise ValueError(f"{name} guide must have {channels} channel(s)")
    return arr


def denoise(linear, radius=1, sigma_color=0.08, amount=1.0, *, normal=None, albedo=None, depth=None,
            sigma_normal=0.35, sigma_albedo=0.20, sigma_depth=1.0, dilation=1):
    if radius <= 0 or amount <= 0.0:
        return linear
    src = np.asarray(linear, dtype=np.float32)
    r = int(radius)
    d = int(dilation)
    pad = r * d
    h, w = src.shape[:2]
    amount = min(max(float(amount), 0.0), 1.0)
    two_s2 = 2.0 * max(float(sigma_color) ** 2, 1e-8)

    def padded(a):
        return np.pad(a, ((pad, pad), (pad, pad), (0, 0)), mode="reflect")

    psrc = padded(src)
    guides = []
    for name, g, sg, ch in (("normal", normal, sigma_normal, 3), ("albedo", albedo, sigma_albedo, 3),
                            ("depth", depth, sigma_depth, 1)):
        g = _guide(name, g, (h, w), ch)
        if g is not None:
            guides.append((padded(g), g, 2.0 * max(float(sg) ** 2, 1e-8)))

    acc = np.zeros_like(src)
    wsum = np.zeros((h, w), np.float32)
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            oy, ox = pad + dy * d, pad + dx * d
            q = psrc[oy:oy + h, ox:ox + w]
            e = q - src
            wgt = (1.0 / (1.0 + dx * dx + dy * dy)) * np.exp(-np.sum(e * e, axis=2) / two_s2).astype(np.float32)
            for pg, g, ts in guides:
                ge = pg[oy:oy + h, ox:ox + w] - g
                wgt *= np.exp(-np.sum(ge * ge, axis=2) / ts).astype(np.float32)
            acc += q * wgt[:, :, None]
            wsum += wgt
    filt = acc / np.maximum(wsum[:, :, None], 1e-8)
    return src * (1.0 - amount) + filt * amount


def display(linear):
    """rendering/denoise.py:107-108"""
    return np.sqrt(np.minimum(np.maximum(linear, 0.0), 1.0)).astype(np.float32)
