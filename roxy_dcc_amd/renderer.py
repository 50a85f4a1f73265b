"""``TaichiRenderer`` drop-in: rendering/taichi_renderer.py:23-203 on gfx950.

Same constructor signature, defaults and validation (``ValueError`` for a
bad ``direct_light_mode``, ``direct_light_max_depth=None`` meaning
``max_depth``), same ``render()`` contract (writes display values into
``image.pixels``), same ``last_ray_count`` / ``last_stats`` / ``__repr__``.

Differences by design: the scene is flattened once and its BVH built and
kept in HBM; samples are traced as wavefront passes of many samples each
(``samples_per_pass``) instead of one launch per sample; ray counting is
always on (wave-aggregated counters), and the denoise runs on the device.
Sampling is Philox keyed by ``seed``, so renders are reproducible.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _native as nat
from .extract import extract_scene
from .render_stats import RenderStats

# Paths traced together per wavefront pass when samples_per_pass is not
# given.  Every traversal launch ends with a tail in which the last deep
# rays walk the BVH one dependent node fetch at a time (about 0.3-0.5 ms on
# the scenes measured); a pass of 2^25 paths (about 200 bytes of workspace
# each, 6.7 GB) spreads that tail over enough rays that it costs a few per
# cent.  With a viewport attached, passes stay near 2^21 paths so the
# display refreshes every few tens of milliseconds.
TARGET_PATHS_PER_PASS = 1 << 25
VIEWPORT_PATHS_PER_PASS = 1 << 21


def default_samples_per_pass(width, height, samples, paths=TARGET_PATHS_PER_PASS):
    per = max(1, paths // max(1, width * height))
    return int(max(1, min(samples, per)))


class DeviceFrame:
    """Per-pixel running means and the display buffer, resident in HBM."""

    def __init__(self, width, height, device="cuda"):
        import torch

        self.width, self.height = int(width), int(height)
        n = self.width * self.height
        self.color = torch.zeros((self.height, self.width, 3), dtype=torch.float32, device=device)
        self.normal = torch.zeros_like(self.color)
        self.albedo = torch.zeros_like(self.color)
        self.depth = torch.zeros((self.height, self.width), dtype=torch.float32, device=device)
        self.display = torch.zeros_like(self.color)
        self.samples = 0
        assert self.color.numel() == 3 * n

    def struct(self):
        return nat.RxAccum(self.color.data_ptr(), self.normal.data_ptr(), self.albedo.data_ptr(),
                           self.depth.data_ptr(), self.display.data_ptr())


class Integrator:
    """Wavefront path tracer over one DeviceScene (``rx_render``).

    ``eye_copies``: size the workspace for the scene's eye-relative copies
    (``rx_render_workspace_size_for_scene``), which the camera-ray traversal
    reads instead of recomputing box - eye, eye - v0 and the primitive
    terms per ray; False renders the same images without them."""

    def __init__(self, scene, width, height, samples_per_pass, device="cuda", eye_copies=True):
        import torch

        self.scene = scene
        self.width, self.height = int(width), int(height)
        self.samples_per_pass = int(samples_per_pass)
        ws = C.c_size_t()
        if eye_copies:
            nat.check(nat.lib().rx_render_workspace_size_for_scene(C.byref(scene.struct()), self.width, self.height,
                                                                   self.samples_per_pass, C.byref(ws)),
                      "render workspace")
        else:
            nat.check(nat.lib().rx_render_workspace_size(self.width, self.height, self.samples_per_pass,
                                                         scene.builder.depth, C.byref(ws)), "render workspace")
        self.workspace = torch.empty(ws.value, dtype=torch.uint8, device=device)
        self.counters = torch.zeros(8, dtype=torch.int64, device=device)
        self.profile = nat.RxProfile()

    def render(self, frame, camera, params, first_sample, n_samples, stream=None):
        st = nat.lib().rx_render(C.byref(self.scene.struct()), C.byref(camera), C.byref(params), int(first_sample),
                                 int(n_samples), self.workspace.data_ptr(), self.workspace.numel(),
                                 C.byref(frame.struct()), self.counters.data_ptr(), C.byref(self.profile),
                                 nat.stream_ptr(stream))
        nat.check(st, "render")
        frame.samples = first_sample + n_samples

    def counts(self):
        v = self.counters.cpu().numpy().astype(np.uint64)
        names = ("rays", "shadow_rays", "node_pops", "tri_tests", "prim_tests", "shadow_node_pops",
                 "shadow_tri_tests", "shadow_prim_tests")
        return {k: int(x) for k, x in zip(names, v)}

    def reset_counts(self):
        self.counters.zero_()
        self.profile = nat.RxProfile()


def camera_struct(cam, width, height):
    """render_kernel's camera arguments (taichi_renderer.py:92-98)."""
    fov_tan = math.tan(math.radians(cam.fov) / 2)
    return nat.RxCamera((C.c_float * 3)(*[float(x) for x in cam.position]),
                        (C.c_float * 3)(*[float(x) for x in cam.forward]),
                        (C.c_float * 3)(*[float(x) for x in cam.right]),
                        (C.c_float * 3)(*[float(x) for x in cam.up]),
                        fov_tan, float(cam.aspect_ratio), int(width), int(height))


def params_struct(max_depth, use_sky, mode_id, dl_max_depth, sample_clamp, bg, seed, samples_per_pass, flags=0,
                  trav_blocks_per_cu=0, adaptive=None):
    """adaptive: None, or (threshold, min_samples, check_interval) to set
    RX_FLAG_ADAPTIVE."""
    seed = int(seed) & ((1 << 64) - 1)
    thr, amin, aint = 0.0, 1, 1
    if adaptive is not None:
        thr, amin, aint = adaptive
        flags |= nat.RX_FLAG_ADAPTIVE
    return nat.RxRenderParams(int(max_depth), int(use_sky), int(mode_id), int(dl_max_depth), float(sample_clamp),
                              (C.c_float * 3)(*[float(x) for x in bg]), seed & 0xFFFFFFFF, seed >> 32, int(flags),
                              int(samples_per_pass), int(trav_blocks_per_cu), float(thr), int(amin), int(aint))


class TaichiRenderer:
    """GPU path tracer for one MI355X behind the reference renderer interface."""

    def __init__(self, world, image, viewport, samples=16, max_depth=4, direct_light_mode="one", denoise=False,
                 denoise_radius=1, denoise_sigma=0.08, denoise_amount=0.8, sample_clamp=10.0,
                 direct_light_max_depth=1, startup_progress=None, *, seed=0, samples_per_pass=None,
                 denoise_passes=1, adaptive_sampling=False, adaptive_min_samples=4, adaptive_threshold=0.002,
                 adaptive_check_interval=1, device="cuda", verbose=True):
        # adaptive stop as the CPU tracer has it (rendering/ray_tracer.py:292-323)
        if adaptive_min_samples < 1:
            raise ValueError("adaptive_min_samples must be at least 1")
        if adaptive_threshold < 0:
            raise ValueError("adaptive_threshold must be non-negative")
        if adaptive_check_interval < 1:
            raise ValueError("adaptive_check_interval must be at least 1")
        self._adaptive_sampling = bool(adaptive_sampling)
        self._adaptive_min_samples = int(adaptive_min_samples)
        self._adaptive_threshold = float(adaptive_threshold)
        self._adaptive_check_interval = int(adaptive_check_interval)
        self._adaptive_stopped = False
        self._last_adaptive_error = None
        self._world = world
        self._image = image
        self._viewport = viewport
        self._samples = samples
        self._max_depth = max_depth
        self._direct_light_mode = direct_light_mode
        self._direct_light_mode_id = self._direct_light_mode_to_id(direct_light_mode)
        if direct_light_max_depth is None:
            direct_light_max_depth = max_depth
        self._direct_light_max_depth = max(0, int(direct_light_max_depth))
        self._denoise = denoise
        self._denoise_radius = denoise_radius
        self._denoise_sigma = denoise_sigma
        self._denoise_amount = denoise_amount
        self._denoise_passes = int(denoise_passes)
        self._sample_clamp = sample_clamp
        self._camera = world.active_camera
        self._last_ray_count = 0
        self._last_stats = None
        self._startup_progress = startup_progress
        self._seed = seed
        self._samples_per_pass = samples_per_pass
        self._device = device
        self._verbose = verbose
        self.scene = None
        self.frame = None
        self.integrator = None

    @staticmethod
    def _direct_light_mode_to_id(mode):
        if mode in (0, "one", "random", "sample"):
            return 0
        if mode in (1, "all", "final"):
            return 1
        raise ValueError("direct_light_mode must be 'one' or 'all'")

    def _denoised_display(self):
        from .denoise import edge_aware_denoise, linear_to_display

        f = self.frame
        x = f.color
        for k in range(max(1, self._denoise_passes)):
            x = edge_aware_denoise(x, radius=self._denoise_radius, sigma_color=self._denoise_sigma,
                                   amount=self._denoise_amount, normal=f.normal, albedo=f.albedo, depth=f.depth,
                                   dilation=1 << k, device=self._device)
        return linear_to_display(x)

    def _copy_display_pixels(self, apply_denoise=False):
        import torch

        if apply_denoise and self._denoise:
            out = self._denoised_display()
        else:
            out = self.frame.display
        px = self._image.pixels
        if (isinstance(px, np.ndarray) and px.dtype == np.float32 and px.flags.c_contiguous
                and px.shape == tuple(out.shape) and out.is_contiguous()):
            # straight into the image's own buffer: a fresh host array for a
            # 1080p frame costs about 4.5 ms of first-touch page faults
            torch.from_numpy(px).copy_(out)
        else:
            px[:] = out.cpu().numpy()

    def render(self):
        import torch

        nat.require_gpu()
        W, H = self._image.width, self._image.height
        cam = self._camera
        world = self._world
        use_sky = int(world._use_sky)
        bg = list(world._background_color)
        t0 = time.perf_counter()
        if self._startup_progress:
            self._startup_progress.step("Extracting scene for GPU", "Flattening objects, materials and lights.")
        stats = extract_scene(world, device=self._device)
        self.scene = stats.scene
        spp = self._samples_per_pass or default_samples_per_pass(
            W, H, self._samples, VIEWPORT_PATHS_PER_PASS if self._viewport else TARGET_PATHS_PER_PASS)
        if self._startup_progress:
            self._startup_progress.close()
            self._startup_progress = None
        self._adaptive_stopped = False
        self._last_adaptive_error = None
        if self._adaptive_sampling:
            done, rays = self._render_adaptive(W, H, spp, use_sky, bg)
        else:
            done, rays = self._render_fixed(W, H, spp, use_sky, bg)
        torch.cuda.synchronize()
        self._copy_display_pixels(apply_denoise=True)
        self._last_ray_count = rays
        s = self.scene.stats()
        self._last_stats = RenderStats(width=W, height=H, samples_requested=self._samples, samples_rendered=done,
                                       max_depth=self._max_depth, rays_cast=rays,
                                       elapsed_seconds=time.perf_counter() - t0,
                                       primitive_count=s["primitive_count"], bvh_nodes=s["bvh_nodes"],
                                       bvh_triangles=s["bvh_triangles"], bvh_materials=s["bvh_materials"],
                                       bvh_leaf_size=s["bvh_leaf_size"], adaptive_sampling=self._adaptive_sampling,
                                       adaptive_stopped=self._adaptive_stopped,
                                       adaptive_threshold=self._adaptive_threshold,
                                       adaptive_error=self._last_adaptive_error,
                                       adaptive_min_samples=self._adaptive_min_samples)
        if self._verbose:
            print("\n" + self._last_stats.format_report())

    def _viewport_tick(self):
        if self._viewport:
            self._copy_display_pixels()
            self._viewport.update(self._image)
            self._viewport.poll_events()

    def _closing(self):
        return bool(self._viewport and getattr(self._viewport, "should_close", False))

    def _render_fixed(self, W, H, spp, use_sky, bg):
        """All requested samples in passes of `spp` (rx_render)."""
        self.frame = DeviceFrame(W, H, device=self._device)
        self.integrator = Integrator(self.scene, W, H, spp, device=self._device)
        camera = camera_struct(self._camera, W, H)
        params = params_struct(self._max_depth, use_sky, self._direct_light_mode_id, self._direct_light_max_depth,
                               self._sample_clamp, bg, self._seed, spp)
        done = 0
        while done < self._samples and not self._closing():
            n = min(spp, self._samples - done)
            self.integrator.render(self.frame, camera, params, done, n)
            done += n
            self._viewport_tick()
        return done, self.integrator.counts()["rays"]

    def _render_adaptive(self, W, H, spp, use_sky, bg):
        """The adaptive stop of the CPU tracer (rendering/ray_tracer.py:531-541)
        decided on the device after every sample: a one-job batch whose record
        reports the samples taken, the rays they cast and the last error."""
        from .batch import Batch, RenderJob

        job = RenderJob(self.scene, self._camera, W, H, self._samples, max_depth=self._max_depth,
                        direct_light_mode=self._direct_light_mode, direct_light_max_depth=self._direct_light_max_depth,
                        sample_clamp=self._sample_clamp, use_sky=bool(use_sky), background=tuple(bg),
                        seed=self._seed, samples_per_pass=spp,
                        adaptive=(self._adaptive_threshold, self._adaptive_min_samples,
                                  self._adaptive_check_interval))
        job.frame = self.frame = DeviceFrame(W, H, device=self._device)
        self.integrator = None
        batch = Batch([job], use_graph=False, device=self._device)
        rec = None
        for _ in range(job.passes_needed):
            if self._closing():
                break
            batch.run(1)
            rec = batch.read_jobs()[0]
            self._viewport_tick()
            if rec.samples_left <= 0:
                break
        batch.close()
        if rec is None:
            return 0, 0
        self._adaptive_stopped = bool(rec.adaptive_stopped)
        self._last_adaptive_error = None if rec.adaptive_error < 0 else float(rec.adaptive_error)
        # the job's rays are those of the samples it kept (a pass's samples
        # past the stop are traced but neither accumulated nor counted)
        return int(rec.first_sample), int(rec.rays)

    @property
    def last_ray_count(self):
        return self._last_ray_count

    @property
    def last_stats(self):
        return self._last_stats

    def __repr__(self):
        return (f"TaichiRenderer(samples={self._samples}, max_depth={self._max_depth}, "
                f"direct_light_mode={self._direct_light_mode!r})")


HipRenderer = TaichiRenderer
