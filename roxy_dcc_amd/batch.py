"""Many scenes in one wavefront: the batched form of the render pass.

Each ``RenderJob`` is one scene (its own resident BVH), camera, resolution,
sample target and accumulation frame.  ``Batch`` hands the job records to
``rx_batch_create`` once; every ``run(passes)`` then traces a pass of all
unfinished jobs together (one set of kernel launches for the whole batch,
replayed from a HIP graph).  Because every random draw is keyed by the
job's seed and (pixel, sample, bounce), a job's image is bit-identical
whether it is rendered alone, in a batch, or split across passes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

from . import _native as nat
from .renderer import DeviceFrame, TaichiRenderer, camera_struct, params_struct


@dataclass
class RenderJob:
    scene: object            # DeviceScene
    camera: object           # world camera (position/forward/right/up/fov/aspect_ratio)
    width: int
    height: int
    samples: int
    max_depth: int = 4
    direct_light_mode: str = "one"
    direct_light_max_depth: int = 1
    sample_clamp: float = 10.0
    use_sky: bool = False
    background: tuple = (0.0, 0.0, 0.0)
    seed: int = 0
    samples_per_pass: int = 1
    trav_blocks_per_cu: int = 0     # launch shape (see autotune.Tuning)
    flags: int = 0
    # adaptive stop (rendering/ray_tracer.py:531-541): None, or
    # (threshold, min_samples, check_interval)
    adaptive: tuple = None
    first_sample: int = 0
    frame: DeviceFrame = field(default=None, repr=False)
    tag: object = None

    def record(self, flags=0) -> nat.RxJob:
        if self.frame is None:
            self.frame = DeviceFrame(self.width, self.height, device=self.scene.device)
        mode = TaichiRenderer._direct_light_mode_to_id(self.direct_light_mode)
        j = nat.RxJob()
        j.scene = self.scene.struct()
        j.camera = camera_struct(self.camera, self.width, self.height)
        j.params = params_struct(self.max_depth, int(self.use_sky), mode, self.direct_light_max_depth,
                                 self.sample_clamp, self.background, self.seed, self.samples_per_pass,
                                 flags | self.flags, self.trav_blocks_per_cu, self.adaptive)
        j.accum = self.frame.struct()
        j.first_sample = self.first_sample
        j.samples_left = self.samples - self.first_sample
        j.rays = 0
        j.adaptive_error = -1.0
        j.adaptive_stopped = 0
        return j

    @property
    def passes_needed(self):
        left = self.samples - self.first_sample
        return (left + self.samples_per_pass - 1) // self.samples_per_pass


class Batch:
    """A set of jobs traced together (``rx_batch_*``)."""

    def __init__(self, jobs, use_graph=True, flags=0, device="cuda"):
        import torch

        nat.require_gpu()
        self.jobs = list(jobs)
        n = len(self.jobs)
        self._records = (nat.RxJob * n)(*[j.record(flags) for j in self.jobs])
        ws = C.c_size_t()
        nat.check(nat.lib().rx_batch_workspace_size(self._records, n, C.byref(ws)), "batch workspace")
        self.workspace = torch.empty(ws.value, dtype=torch.uint8, device=device)
        self.counters = torch.zeros(8, dtype=torch.int64, device=device)
        h = C.c_void_p()
        nat.check(nat.lib().rx_batch_create(self._records, n, self.workspace.data_ptr(), ws.value, int(use_graph),
                                            C.byref(h)), "batch create")
        self._h = h
        self.passes_run = 0

    @property
    def passes_needed(self):
        return max(j.passes_needed for j in self.jobs)

    def run(self, passes=None, profile=None, stream=None):
        passes = self.passes_needed - self.passes_run if passes is None else passes
        if passes <= 0:
            return
        nat.check(nat.lib().rx_batch_run(self._h, int(passes), self.counters.data_ptr(),
                                         C.byref(profile) if profile is not None else None,
                                         nat.stream_ptr(stream)), "batch run")
        self.passes_run += passes

    def graph_nodes(self):
        """(kernel, memset, other) node counts of the captured pass, or
        (-1, -1, -1) before a graph has been captured."""
        k, m, o = C.c_int(), C.c_int(), C.c_int()
        nat.check(nat.lib().rx_batch_graph_nodes(self._h, C.byref(k), C.byref(m), C.byref(o)), "batch graph nodes")
        return k.value, m.value, o.value

    def read_jobs(self, stream=None):
        out = (nat.RxJob * len(self.jobs))()
        nat.check(nat.lib().rx_batch_read_jobs(self._h, out, nat.stream_ptr(stream)), "batch read")
        return list(out)

    def rays(self):
        return int(self.counters[0].item())

    def close(self):
        if getattr(self, "_h", None):
            nat.lib().rx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_alone(job: RenderJob, use_graph=False):
    """Render one job by itself through the same batch path."""
    b = Batch([job], use_graph=use_graph)
    b.run()
    b.close()
    return job.frame
