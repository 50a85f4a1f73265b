"""Edge-aware denoise on the GPU (rendering/denoise.py:23-108 interface).

``edge_aware_denoise`` keeps the reference signature, defaults, guide
validation and disable rules; the filtering itself runs in ``rx_denoise``
(LDS-tiled kernel).  Inputs may be numpy arrays (copied to the device and
the result copied back, as the reference returns numpy) or torch device
tensors (kept on the device).  ``dilation`` spaces the taps for a-trous
passes; ``atrous`` runs the dilation 1, 2, 4, ... sequence.
"""
from __future__ import annotations

import numpy as np

from . import _native as nat


def _as_device(x, torch, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float32).contiguous(), True
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float32))).to(device), False


def _guide(name, guide, shape, channels, torch, device):
    if guide is None:
        return None
    t, _ = _as_device(guide, torch, device)
    if tuple(t.shape[:2]) != tuple(shape):
        raise ValueError(f"{name} guide must match pixel height and width")
    if t.dim() == 2:
        t = t[:, :, None]
    elif t.dim() != 3:
        raise ValueError(f"{name} guide must be a 2D or 3D array")
    if t.shape[2] != channels:
        raise ValueError(f"{name} guide must have {channels} channel(s)")
    return t.contiguous()


def edge_aware_denoise(linear_pixels, radius=1, sigma_color=0.08, amount=1.0, *, normal=None, albedo=None,
                       depth=None, sigma_normal=0.35, sigma_albedo=0.20, sigma_depth=1.0, dilation=1,
                       device="cuda", stream=None):
    """Guided bilateral filter; see rendering/denoise.py:23-40."""
    if radius <= 0 or amount <= 0.0:
        return linear_pixels
    import torch

    nat.require_gpu()
    src, on_device = _as_device(linear_pixels, torch, device)
    if src.dim() != 3 or src.shape[2] != 3:
        raise ValueError("linear_pixels must be an (H, W, 3) array")
    h, w = int(src.shape[0]), int(src.shape[1])
    g_n = _guide("normal", normal, (h, w), 3, torch, device)
    g_a = _guide("albedo", albedo, (h, w), 3, torch, device)
    g_d = _guide("depth", depth, (h, w), 1, torch, device)
    out = torch.empty_like(src)
    amt = min(max(float(amount), 0.0), 1.0)
    nat.check(nat.lib().rx_denoise(
        src.data_ptr(), nat.ptr(g_n), nat.ptr(g_a), nat.ptr(g_d), w, h, int(radius), int(dilation),
        float(sigma_color), float(sigma_normal), float(sigma_albedo), float(sigma_depth), amt, out.data_ptr(),
        nat.stream_ptr(stream)), "denoise")
    if on_device:
        return out
    return out.cpu().numpy()


def atrous(linear_pixels, passes=5, radius=1, sigma_color=0.08, amount=1.0, **guides):
    """``passes`` filter passes with dilations 1, 2, 4, ...; each pass
    filters the previous pass's output with the same guides."""
    x = linear_pixels
    for k in range(passes):
        x = edge_aware_denoise(x, radius=radius, sigma_color=sigma_color, amount=amount, dilation=1 << k, **guides)
    return x


def linear_to_display(linear_pixels, device="cuda", stream=None):
    """sqrt(clip(x, 0, 1)) (rendering/denoise.py:107-108)."""
    import torch

    if not isinstance(linear_pixels, torch.Tensor):
        nat.require_gpu()
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(linear_pixels, dtype=np.float32))).to(device)
        out = torch.empty_like(t)
        nat.check(nat.lib().rx_display(t.data_ptr(), t.numel(), out.data_ptr(), nat.stream_ptr(stream)), "display")
        return out.cpu().numpy()
    t = linear_pixels.contiguous()
    out = torch.empty_like(t)
    nat.check(nat.lib().rx_display(t.data_ptr(), t.numel(), out.data_ptr(), nat.stream_ptr(stream)), "display")
    return out
