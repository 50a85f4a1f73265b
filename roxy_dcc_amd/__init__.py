"""MI355X (gfx950) path-tracing backend behind the Roxy ``TaichiRenderer`` interface.

Host code in Python drives hand-written HIP kernels through the C ABI of
``librxhip.so`` (include/roxy_hip.h); torch-ROCm only provides device
buffers and streams.
"""
from .bvh import GPUBVHBuilder, TriangleBatch
from .denoise import atrous, edge_aware_denoise, linear_to_display
from .extract import extract_scene, flatten_scene
from .render_stats import RenderStats
from .renderer import TaichiRenderer, HipRenderer

__all__ = [
    "GPUBVHBuilder", "TriangleBatch", "edge_aware_denoise", "linear_to_display", "atrous", "extract_scene",
    "flatten_scene", "RenderStats", "TaichiRenderer", "HipRenderer",
]
