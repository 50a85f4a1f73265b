"""Continuous-batching render server.

Requests (scene, resolution, samples) arrive over time.  The server keeps
up to ``max_scenes`` scene BVHs resident (least recently used evicted,
never one an active job needs), admits waiting requests in arrival order
while job slots and the frame-memory budget allow, and every iteration
traces one wavefront pass of the active jobs together (``Batch``).  The
pass's path budget is shared out by ``share_pass`` (shortest remaining work
first, guarded by first-come-first-served deadlines), so the device stays
full while small jobs finish in a pass or two and large ones take what is
left (a large job shares a pass only when enough jobs queue behind it).
Finished jobs are retired and their frames read back while the next pass
traces; new arrivals join at the next iteration, so their work overlaps
with jobs already on the GPU.  A job's image is the same alone and in any mix
(sampling is keyed by job seed, pixel, sample and bounce; accumulation is
per job in sample order), whatever samples per pass it is given.

``serial_loop`` is the one-job-at-a-time baseline: each request is rendered
to completion, in arrival order, before the next starts.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from .batch import Batch, RenderJob
from .renderer import TARGET_PATHS_PER_PASS


@dataclass
class Request:
    rid: int
    scene_key: object
    width: int
    height: int
    samples: int
    arrival: float            # seconds after the server clock starts
    max_depth: int = 6
    seed: int = 0
    # error target instead of a fixed sample count: (threshold, min_samples,
    # check_interval) of the adaptive stop (rendering/ray_tracer.py:531-541),
    # `samples` then being the cap
    adaptive: tuple = None
    # render settings; None takes the scene's own (the world's sky or
    # background and NEE settings, when the cache's loader gives them)
    use_sky: bool = None
    background: tuple = None
    direct_light_mode: str = None
    direct_light_max_depth: int = None
    # filled in by the server
    admitted: float = None
    finished: float = None
    samples_done: int = 0
    frame: object = field(default=None, repr=False)
    rays: int = 0

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class SceneCache:
    """Resident DeviceScenes keyed by scene key, LRU under a byte budget.
    The loader maps a key to (DeviceScene, camera) or (DeviceScene, camera,
    settings), settings a dict of the scene's own render settings (use_sky,
    background, direct_light_mode, direct_light_max_depth); ``get`` returns
    the triple."""

    def __init__(self, loader, max_scenes=32, max_bytes=64 << 30):
        self.loader = loader
        self.max_scenes = max_scenes
        self.max_bytes = max_bytes
        self.entries = collections.OrderedDict()
        self.loads = 0
        self.evictions = 0

    def bytes(self):
        return sum(e[0].hbm_bytes() for e in self.entries.values())

    def get(self, key, pinned=()):
        if key in self.entries:
            self.entries.move_to_end(key)
            return self.entries[key]
        got = tuple(self.loader(key))
        entry = got if len(got) == 3 else (got[0], got[1], {})
        self.loads += 1
        self.entries[key] = entry
        self._evict(pinned=set(pinned) | {key})
        return entry

    def _evict(self, pinned):
        while (len(self.entries) > self.max_scenes or self.bytes() > self.max_bytes):
            victim = next((k for k in self.entries if k not in pinned), None)
            if victim is None:
                break
            del self.entries[victim]
            self.evictions += 1


def fifo_deadlines(work, slack):
    """Deadlines of jobs served one after another in arrival order: job k
    finishes after work[0] + ... + work[k], plus ``slack``."""
    out, run = [], 0
    for w in work:
        run += w
        out.append(run + slack)
    return out


def deadline_order(work, deadline):
    """Shortest remaining work first, guarded by deadlines.  ``work`` is
    each job's remaining work and ``deadline`` the work done from now by
    which it must finish, both in the same unit.  The order starts as
    shortest-first; while some job would finish past its deadline, the
    first such job moves ahead of the first job with a later deadline.  If
    the moves have not converged after n^2 of them, the order is earliest
    deadline first, which meets every deadline any order can meet."""
    n = len(work)
    order = sorted(range(n), key=lambda k: (work[k], k))
    for _ in range(n * n + 1):
        t, late = 0, None
        for pos, k in enumerate(order):
            t += work[k]
            if t > deadline[k]:
                late = pos
                break
        if late is None:
            return order
        k = order.pop(late)
        ins = next((i for i, j in enumerate(order) if deadline[j] > deadline[k]), late)
        order.insert(ins, k)
    return sorted(range(n), key=lambda k: (deadline[k], k))


def share_pass(demands, budget, policy="deadline", deadlines=None, costs=None, join=None):
    """Split of a pass's path budget: `demands` are (pixels, samples_left)
    per job in arrival order; returns samples per pass for each (whole
    samples, at most what the job still needs; 0 = not in this pass).

    "deadline" (the default): jobs take the budget in deadline_order
    (shortest remaining work first unless that would make some job finish
    past its deadline), each as many samples as it can use; jobs whose
    next sample no longer fits wait for a later pass, and the first job in
    the order always advances; a job whose remaining samples overshoot the
    budget by at most a quarter of it takes them all.  With ``join`` (a
    pass's fixed cost, in paths traced in the same time), a job joins a
    pass that already holds others only while (jobs in the pass) x (the
    paths it takes) is at most ``join`` x (it and the jobs after it in the
    order): joining delays every job of the pass by its paths, waiting
    delays it and the jobs behind it by one pass's fixed cost; the first
    job that does not join ends the pass.  A job's work is
    pixels x samples left x its
    cost per path (``costs``, 1 by default); ``deadlines`` are in the same
    unit, counted from now (by default: first come first served from now,
    with one pass of slack).  Low median latency like shortest-first, a
    tail close to the arrival-order loop's.
    "fair": water-filling, smallest remaining work first, each job taking
    at most an equal share of what is left, so small jobs finish in one
    pass and the budget they leave goes to the larger ones (processor
    sharing).
    "fifo": jobs take the budget in arrival order, as much as each can use,
    and every other job still advances one sample per pass (first come
    first served with backfill).
    "srpt": as "fifo", but in order of remaining work (pixels x samples
    left), smallest first: shortest remaining processing time with
    backfill; it starves the longest jobs under load."""
    n = len(demands)
    if policy == "deadline":
        out = [0] * n
        costs = costs or [1.0] * n
        work = [px * need * c for (px, need), c in zip(demands, costs)]
        stretch = 0.25
        order = deadline_order(work, fifo_deadlines(work, budget) if deadlines is None else deadlines)
        left = budget
        joined = 0
        for rank, k in enumerate(order):
            px, need = demands[k]
            take = max(0, min(need, left // max(1, px)))
            if rank == 0:
                take = max(1, take)
            # a job whose remainder overshoots the budget by a little finishes
            # in this pass rather than taking a whole pass for its last
            # samples (the pass grows by at most a quarter)
            over = (need - take) * px if take < need else 0
            if over and over <= stretch * budget:
                take = need
            # joining a pass puts off the end of every job already in it by
            # the joiner's paths; waiting for the next pass puts off the
            # joiner and every job after it by one pass's fixed cost (``join``
            # paths' worth of time): join while that costs less in total
            if join is not None and joined and take and joined * take * px > join * (n - rank):
                break
            if take == need and over:
                stretch -= over / budget
            out[k] = take
            left -= take * px
            joined += take > 0
        return out
    out = [1] * n
    if policy in ("fifo", "srpt"):
        order = range(n) if policy == "fifo" else sorted(range(n), key=lambda k: demands[k][0] * demands[k][1])
        left = budget - sum(px for px, _ in demands)
        for k in order:
            px, need = demands[k]
            extra = max(0, min(need - 1, left // max(1, px)))
            out[k] = 1 + extra
            left -= extra * px
        return out
    if policy != "fair":
        raise ValueError(f"unknown pass-sharing policy {policy!r}")
    order = sorted(range(n), key=lambda k: demands[k][0] * demands[k][1])
    left = budget
    for rank, k in enumerate(order):
        px, need = demands[k]
        share = left // (n - rank)
        spp = max(1, min(need, share // max(1, px)))
        out[k] = spp
        left = max(0, left - spp * px)
    return out


def request_job(req: Request, scene, cam, opts, samples_per_pass):
    """The RenderJob of a request: its own settings where given, else the
    scene's (``opts`` from the cache's loader), else RenderJob's defaults."""

    def pick(name, default):
        v = getattr(req, name)
        return opts.get(name, default) if v is None else v

    return RenderJob(scene, cam, req.width, req.height, req.samples, max_depth=req.max_depth, seed=req.seed,
                     samples_per_pass=samples_per_pass, adaptive=req.adaptive, tag=req,
                     use_sky=bool(pick("use_sky", False)), background=tuple(pick("background", (0.0, 0.0, 0.0))),
                     direct_light_mode=pick("direct_light_mode", "one"),
                     direct_light_max_depth=int(pick("direct_light_max_depth", 1)))


class RenderServer:
    """Continuous batching: up to ``max_jobs`` requests share every pass.

    Each iteration traces one wavefront pass of all active jobs together.
    The pass's path budget (``pass_paths``) is split over the active jobs
    by ``share_pass``.  The drop-in renderer's pass size (2^25 paths) keeps
    the traversal kernels at their large-pass rate, so the largest jobs are
    traced as fast as the one-request-at-a-time loop traces them, while a
    pass stays short enough (tens of ms) that new arrivals join soon.  The
    default policy, shortest remaining work first guarded by
    first-come-first-served deadlines, keeps the median latency of
    shortest-first without letting the longest jobs finish much later than
    the arrival-order loop would.  Every job of a pass finishes when the
    pass ends, so a large job joins a pass only when enough jobs queue
    behind it to be worth the delay it adds to the others (``join_paths``);
    small jobs ride along.  Frames of finished jobs are copied to the host
    while the next pass traces, into host arrays touched while their last
    pass traced.  Requests are admitted in arrival order
    while job slots and the frame-memory budget allow."""

    def __init__(self, cache: SceneCache, max_jobs=32, pass_paths=TARGET_PATHS_PER_PASS, frame_bytes_budget=32 << 30,
                 policy="deadline", slack_passes=0.1, join_paths=1 << 22):
        self.cache = cache
        self.policy = policy
        self.join_paths = join_paths
        self.max_jobs = max_jobs
        self.pass_paths = pass_paths
        # how far past its first-come-first-served finish a job may be put
        # off for shorter ones, in passes (a pass cannot end exactly on a
        # deadline)
        self.slack_passes = slack_passes
        # a pass's fixed cost (launch tails, read-back, set-up: 2-3 ms) as
        # paths traced in the same time; None lets every job that fits join
        # a pass (share_pass's ``join``)
        self.frame_bytes_budget = frame_bytes_budget
        self.waiting = collections.deque()
        self.active = []       # (Request, RenderJob)
        self.done = []
        self.passes = 0
        # deadlines, in rays: the rays traced so far, the finish of the last
        # arrival in a first-come-first-served schedule, rays per path by
        # scene (measured on the passes traced so far)
        self.rays_done = 0
        self.fifo_finish = 0.0
        self.rays_per_path = {}
        self.deadline = {}
        self._batch = None
        self._batch_key = None
        self._batch_rays = {}
        self._copy_stream = None
        self._host = {}        # id(job) -> touched host array for its frame
        self.pass_s = 0.0      # wall time inside run + read-back of the passes

    def _close_batch(self):
        if self._batch is not None:
            self._batch.close()
        self._batch = None
        self._batch_key = None

    def _cost(self, req):
        return self.rays_per_path.get(req.scene_key, 0.5 * req.max_depth + 1.0)

    def _note_arrivals(self, now):
        """Each request, when it arrives, gets the finish time (in rays) that
        serving every request one after another in arrival order would give
        it: its deadline."""
        for req in self.waiting:
            if req.arrival > now:
                break
            if req.rid not in self.deadline:
                work = req.width * req.height * req.samples * self._cost(req)
                self.fifo_finish = max(self.fifo_finish, float(self.rays_done)) + work
                self.deadline[req.rid] = self.fifo_finish

    @staticmethod
    def _frame_bytes(req):
        # colour, normal, albedo, display (3 floats) and depth per pixel
        return req.width * req.height * 13 * 4

    def _admit(self, now):
        used = sum(self._frame_bytes(r) for r, _ in self.active)
        while self.waiting and self.waiting[0].arrival <= now and len(self.active) < self.max_jobs:
            req = self.waiting[0]
            need = self._frame_bytes(req)
            if self.active and used + need > self.frame_bytes_budget:
                break
            self.waiting.popleft()
            pinned = [r.scene_key for r, _ in self.active]
            scene, cam, opts = self.cache.get(req.scene_key, pinned=pinned)
            job = request_job(req, scene, cam, opts, samples_per_pass=1)
            req.admitted = now
            self.active.append((req, job))
            used += need

    def run(self, requests, clock=time.perf_counter, idle_sleep=0.0005):
        """Serve ``requests`` (sorted by arrival); returns the finished list."""
        import torch

        self.waiting.extend(sorted(requests, key=lambda r: r.arrival))
        t0 = clock()
        # finished jobs whose frames are still on the device: they are read
        # back on a side stream while the next pass traces (a host copy of a
        # 1080p frame costs 1.5 ms into touched memory, about 6 ms into a
        # fresh array)
        pending = []
        while self.waiting or self.active or pending:
            now = clock() - t0
            self._note_arrivals(now)
            self._admit(now)
            if not self.active:
                pending = self._read_back(pending, clock, t0)
                if self.waiting:
                    time.sleep(idle_sleep)
                continue
            jobs = [j for _, j in self.active]
            costs = [self._cost(r) for r, _ in self.active]
            slack = self.slack_passes * self.pass_paths * sum(costs) / len(costs)
            deadlines = [self.deadline[r.rid] - self.rays_done + slack for r, _ in self.active]
            spp = share_pass([(j.width * j.height, j.samples - j.first_sample) for j in jobs], self.pass_paths,
                             self.policy, deadlines=deadlines, costs=costs, join=self.join_paths)
            for j, k in zip(jobs, spp):
                j.samples_per_pass = k
            traced = [(r, j) for (r, j), k in zip(self.active, spp) if k > 0]
            # the same jobs with the same shares as the last pass: trace the
            # next pass of the same batch (its records advance on the device)
            key = tuple((r.rid, j.samples_per_pass) for r, j in traced)
            if self._batch is None or key != self._batch_key:
                self._close_batch()
                self._batch = Batch([j for _, j in traced], use_graph=False)
                self._batch_key = key
                self._batch_rays = {}
            batch = self._batch
            tr = clock()
            batch.run(1)
            pending = self._read_back(pending, clock, t0)
            # host arrays for the frames this pass will finish, touched while
            # it traces (a copy into touched memory is about 4x faster than
            # into a fresh array, whose pages fault on first write)
            for (req, job), k in zip(self.active, spp):
                if k and k >= job.samples - job.first_sample and id(job) not in self._host:
                    h = np.empty(tuple(job.frame.color.shape), np.float32)
                    h[...] = 0.0
                    self._host[id(job)] = h
            recs = batch.read_jobs()
            self.pass_s += clock() - tr
            self.passes += 1
            still = [(r, j) for (r, j), k in zip(self.active, spp) if k == 0]
            finished = []
            for (req, job), rec in zip(traced, recs):
                rays = int(rec.rays) - self._batch_rays.get(req.rid, 0)
                self._batch_rays[req.rid] = int(rec.rays)
                traced_paths = job.width * job.height * (int(rec.first_sample) - job.first_sample)
                if traced_paths > 0 and rays > 0:
                    rpp = rays / traced_paths
                    old = self.rays_per_path.get(req.scene_key)
                    self.rays_per_path[req.scene_key] = rpp if old is None else 0.5 * (old + rpp)
                self.rays_done += rays
                job.first_sample = rec.first_sample
                req.rays += rays
                req.samples_done = int(rec.first_sample)
                if rec.samples_left <= 0:
                    finished.append((req, job))
                else:
                    still.append((req, job))
            if finished:
                self._close_batch()
                pending = finished
            self.active = still
        self._close_batch()
        return self.done

    def _read_back(self, finished, clock, t0):
        """Copies the frames of finished jobs to the host, on a stream of its
        own (their passes have completed: ``read_jobs`` waited for them), and
        marks the requests done; returns an empty list."""
        import torch

        if finished:
            if self._copy_stream is None:
                self._copy_stream = torch.cuda.Stream()
            with torch.cuda.stream(self._copy_stream):
                for req, job in finished:
                    h = self._host.pop(id(job), None)
                    if h is None:
                        req.frame = job.frame.color.cpu().numpy()
                    else:
                        torch.from_numpy(h).copy_(job.frame.color)
                        req.frame = h
                    req.finished = clock() - t0
                    self.done.append(req)
        return []


def serial_loop(cache: SceneCache, requests, paths_per_pass=TARGET_PATHS_PER_PASS, clock=time.perf_counter):
    """One request at a time, in arrival order, each rendered to completion
    in passes of up to `paths_per_pass` paths (the drop-in renderer's
    default; 1 = one sample per pass, the reference renderer's
    launch-per-sample model)."""
    import torch

    t0 = clock()
    out = []
    for req in sorted(requests, key=lambda r: r.arrival):
        while clock() - t0 < req.arrival:
            time.sleep(0.0005)
        req.admitted = clock() - t0
        scene, cam, opts = cache.get(req.scene_key)
        spp = int(max(1, min(req.samples, paths_per_pass // max(1, req.width * req.height))))
        job = request_job(req, scene, cam, opts, samples_per_pass=spp)
        b = Batch([job], use_graph=False)
        b.run()
        rec = b.read_jobs()[0]
        b.close()
        torch.cuda.synchronize()
        req.rays = int(rec.rays)
        req.samples_done = int(rec.first_sample)
        req.frame = job.frame.color.cpu().numpy()
        req.finished = clock() - t0
        out.append(req)
    return out


def poisson_arrivals(rate, n, seed=5):
    """n arrival times of a Poisson process with ``rate`` requests/second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def latency_stats(done, wall):
    lat = np.asarray([r.latency for r in done])
    return {"completed": len(done), "wall_s": wall, "renders_per_s": len(done) / wall if wall > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "rays": int(sum(r.rays for r in done))}
