"""Launch-shape autotuner for batched renders.

A batch's images do not depend on how it is scheduled: every random draw is
keyed by the job's seed and (pixel, sample, bounce), and each pixel's
samples are accumulated in sample order.  So the tuner is free to pick,
per job list, whatever runs fastest on the device at hand:

``passes``
    how many wavefront passes the list is split into (each job traces
    ``ceil(samples / passes)`` samples per pass): the queue size of a pass.
``blocks_per_cu``
    resident blocks per CU of the persistent traversal kernels (0 = the
    register budget's maximum).
``small_blocks``
    64-thread traversal blocks with the deep stack and no LDS treelet
    instead of the large blocks (1024 / 512 threads) that share one
    (block shape and stack depth).
``prim_prepass``
    closest-primitive scan as its own pass over each ray queue.

Each candidate is timed on a trial: the first ``trial_passes`` passes of
the real job list with that shape.  The search is coordinate descent (queue
size first, then occupancy, then block shape, then the pre-pass), and the
result is cached by a signature of the job list.
"""
from __future__ import annotations

import dataclasses
import math
import time
from dataclasses import dataclass

from . import _native as nat

# Path slots a pass may hold (int32 path indices) and the default HBM budget
# for a batch workspace (path state is about 200 bytes per slot).
MAX_PATHS_PER_PASS = (1 << 31) - 1
WORKSPACE_BUDGET = 24 << 30
BYTES_PER_PATH = 250


@dataclass(frozen=True)
class Tuning:
    passes: int
    blocks_per_cu: int = 0
    small_blocks: bool = False
    prim_prepass: bool = False

    def flags(self):
        f = 0
        if self.small_blocks:
            f |= nat.RX_FLAG_SMALL_BLOCKS
        if self.prim_prepass:
            f |= nat.RX_FLAG_PRIM_PREPASS
        return f

    def samples_per_pass(self, samples):
        return max(1, math.ceil(samples / self.passes))

    def apply(self, jobs):
        """Sets every job's samples_per_pass and launch shape in place."""
        for j in jobs:
            j.samples_per_pass = self.samples_per_pass(j.samples)
            j.trav_blocks_per_cu = self.blocks_per_cu
            j.flags = self.flags()
        return jobs


def signature(jobs):
    """Cache key: what the launch shape depends on (sizes, depths, scene sizes)."""
    return (len(jobs), sum(j.width * j.height for j in jobs), sum(j.width * j.height * j.samples for j in jobs),
            max(j.max_depth for j in jobs), max(j.scene.builder.depth for j in jobs),
            max(j.scene.n_prims for j in jobs), sum(j.scene.n_tris for j in jobs) // 4096)


def pass_candidates(jobs, budget_bytes=WORKSPACE_BUDGET):
    """Pass counts giving about 2M .. 64M path slots per pass within the budget."""
    total = sum(j.width * j.height * j.samples for j in jobs)
    most = max(j.samples for j in jobs)
    out = []
    for paths in (1 << 21, 1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26):
        k = max(1, min(most, math.ceil(total / paths)))
        per_pass = sum(j.width * j.height * max(1, math.ceil(j.samples / k)) for j in jobs)
        if per_pass > MAX_PATHS_PER_PASS or per_pass * BYTES_PER_PATH > budget_bytes:
            continue
        if k not in out:
            out.append(k)
    if not out:
        out = [most]
    return out


class Autotuner:
    def __init__(self, trial_passes=1, verbose=False):
        self.trial_passes = trial_passes
        self.verbose = verbose
        self.cache = {}
        self.log = []

    def _time(self, jobs, t: Tuning):
        """Samples per second of the first passes of `jobs` with tuning `t`."""
        import torch

        from .batch import Batch

        saved = [(j.samples_per_pass, j.trav_blocks_per_cu, j.flags, j.first_sample) for j in jobs]
        t.apply(jobs)
        for j in jobs:
            j.first_sample = 0
        # warm-up pass on one batch (first touch of the workspace), timed
        # passes on a fresh one over the same memory
        b = Batch(jobs, use_graph=False)
        b.run(1)
        torch.cuda.synchronize()
        b.close()
        b = Batch(jobs, use_graph=False)
        n = min(self.trial_passes, b.passes_needed)
        t0 = time.perf_counter()
        b.run(n)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        b.close()
        samples = sum(j.width * j.height * min(j.samples_per_pass * n, j.samples) for j in jobs)
        for j, (spp, bpc, fl, fs) in zip(jobs, saved):
            j.samples_per_pass, j.trav_blocks_per_cu, j.flags, j.first_sample = spp, bpc, fl, fs
        rate = samples / dt if dt > 0 and samples > 0 else 0.0
        self.log.append((t, rate))
        if self.verbose:
            print(f"  tune {t}: {rate / 1e6:.1f} M samples/s")
        return rate

    def tune(self, jobs):
        key = signature(jobs)
        if key in self.cache:
            return self.cache[key]
        best, best_rate = None, -1.0

        def consider(t):
            nonlocal best, best_rate
            r = self._time(jobs, t)
            if r > best_rate:
                best, best_rate = t, r

        for k in pass_candidates(jobs):
            consider(Tuning(passes=k))
        base = best
        # one block per CU instead of the default two closest-hit / three
        # shadow blocks: half the resident waves, twice the LDS per wave
        consider(dataclasses.replace(base, blocks_per_cu=1))
        if max(j.scene.builder.depth for j in jobs) + 1 <= 24:
            # 64-thread blocks: occupancy back to their own default
            consider(dataclasses.replace(best, small_blocks=True, blocks_per_cu=0))
        if max(j.scene.n_prims for j in jobs) >= 16:
            consider(dataclasses.replace(best, prim_prepass=True))
        self.cache[key] = best
        return best
