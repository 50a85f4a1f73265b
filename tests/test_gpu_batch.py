"""Batched rendering: a job's image is the same alone, in a batch, split
across passes, and replayed from a HIP graph (bit-identical), and equals
the CPU oracle's render of that job."""
from __future__ import annotations

import numpy as np
import pytest

from roxy_dcc_amd import scenes

pytestmark = pytest.mark.gpu


def _jobs(seeds, res_override=None, spp_pass=None):
    from roxy_dcc_amd.batch import RenderJob
    from roxy_dcc_amd.extract import extract_scene

    jobs = []
    for s in seeds:
        world, w, h, st = scenes.batch_scene(s)
        if res_override:
            w = h = res_override
        ex = extract_scene(world)
        jobs.append(RenderJob(ex.scene, world.active_camera, w, h, min(st["samples"], 6), max_depth=st["max_depth"],
                              seed=1000 + s, samples_per_pass=spp_pass or 2, tag=(world, ex)))
    return jobs


def _color(job):
    return job.frame.color.cpu().numpy()


def test_batch_equals_alone_and_graph_equals_direct():
    from roxy_dcc_amd.batch import Batch, render_alone

    import torch

    seeds = [4000, 4001, 4002, 4003]
    mixed = _jobs(seeds, res_override=48)
    b = Batch(mixed, use_graph=True)
    b.run()
    torch.cuda.synchronize()
    recs = b.read_jobs()
    for j, r in zip(mixed, recs):
        assert r.samples_left == 0 and r.first_sample == j.samples and r.rays > 0
    alone = _jobs(seeds, res_override=48, spp_pass=1)
    for a, m in zip(alone, mixed):
        render_alone(a, use_graph=False)
        torch.cuda.synchronize()
        same = np.array_equal(_color(a), _color(m))
        print(f"job {a.seed}: alone==batch {same} mean {_color(a).mean():.4f}", flush=True)
        assert same
    b.close()


def test_reordered_queues_in_batches_change_nothing():
    """Batches trace primaries pixel-major and sort their secondary queues
    too (the sort reads each queue's length on the device, under a HIP
    graph or not): every image equals that of a batch that keeps each queue
    in path order."""
    import torch

    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd.batch import Batch

    seeds = [4020, 4021, 4022]
    imgs = {}
    for name, graph, flags in (("graph", True, 0), ("direct", False, 0),
                               ("kept", False, nat.RX_FLAG_KEEP_QUEUE_ORDER)):
        # 3 jobs x 384^2 x 4 samples per pass: above the pass size (2^20
        # path slots) from which secondary queues are sorted
        jobs = _jobs(seeds, res_override=384, spp_pass=4)
        b = Batch(jobs, use_graph=graph, flags=flags)
        b.run()
        torch.cuda.synchronize()
        imgs[name] = [_color(j) for j in jobs]
        if graph:
            # the captured pass holds the queue sorts (kernels only)
            kernels, memsets, others = b.graph_nodes()
            print(f"captured pass with sorts: {kernels} kernel nodes, {memsets} memset, {others} other", flush=True)
            assert kernels > 40 and memsets == 0 and others == 0
        b.close()
    for g, d, k in zip(imgs["graph"], imgs["direct"], imgs["kept"]):
        print(f"mean {k.mean():.4f}: graph==kept {np.array_equal(g, k)} direct==kept {np.array_equal(d, k)}",
              flush=True)
        assert np.array_equal(g, k) and np.array_equal(d, k)


def test_uneven_jobs_and_oracle_parity():
    """Jobs of different resolutions and sample counts in one batch; each
    matches the oracle's render of that scene."""
    import torch

    from oracle import oracle as orc
    from roxy_dcc_amd.batch import Batch
    from tests import helpers

    jobs = _jobs([4010, 4011, 4012])
    for j, (res, spp) in zip(jobs, [(24, 3), (40, 5), (32, 2)]):
        j.width = j.height = res
        j.samples = spp
    b = Batch(jobs, use_graph=False)
    b.run()
    torch.cuda.synchronize()
    for j in jobs:
        world, ex = j.tag
        osc = helpers.flat_to_oracle(ex.flat)
        ocam = helpers.oracle_camera(world.active_camera, j.width, j.height)
        acc = orc.Accum(j.width * j.height)
        orc.render(osc, ocam, orc.params(j.max_depth, 0, 0, 1, 10.0, (0, 0, 0), j.seed), 0, j.samples, acc)
        got = _color(j).reshape(-1, 3)
        same = np.all(got == acc.color, axis=1).mean()
        print(f"job {j.seed} {j.width}x{j.height}x{j.samples}: exact {same:.4f}", flush=True)
        np.testing.assert_array_equal(got, acc.color)
    b.close()


def test_autotuned_shapes_leave_images_unchanged():
    """Every launch shape the autotuner can pick (queue size, traversal
    occupancy, 64-thread blocks, primitive pre-pass) gives the same images."""
    import dataclasses

    import torch

    from roxy_dcc_amd.autotune import Autotuner, Tuning
    from roxy_dcc_amd.batch import Batch

    seeds = [4020, 4021, 4022]
    ref = _jobs(seeds, res_override=40)
    Batch(ref, use_graph=False).run()
    torch.cuda.synchronize()
    shapes = [Tuning(passes=1), Tuning(passes=3, blocks_per_cu=4), Tuning(passes=2, small_blocks=True),
              Tuning(passes=6, blocks_per_cu=2, prim_prepass=True)]
    for t in shapes:
        jobs = t.apply(_jobs(seeds, res_override=40))
        b = Batch(jobs, use_graph=True)
        b.run()
        torch.cuda.synchronize()
        for r, j in zip(ref, jobs):
            assert np.array_equal(_color(r), _color(j)), (t, j.seed)
        b.close()
    tuner = Autotuner()
    jobs = _jobs(seeds, res_override=40)
    best = tuner.tune(jobs)
    print(f"tuned: {best}; trials {len(tuner.log)}", flush=True)
    assert isinstance(best, Tuning) and len(tuner.log) >= 3
    assert tuner.tune(jobs) is best            # cached by signature
    assert dataclasses.asdict(best)["passes"] >= 1


def test_every_graph_replay_equals_the_eager_pass():
    """The captured pass (which holds the ray-queue radix sort) is replayed
    one pass at a time, with eager device work and synchronisations between
    the replays; after every replay each image and the ray count equal
    those of the same batch run eagerly.  The captured graph holds kernel
    nodes only: no memory-fill node."""
    import torch

    from roxy_dcc_amd.batch import Batch

    seeds = [4030, 4031, 4032]
    graph_jobs = _jobs(seeds, res_override=256, spp_pass=2)
    eager_jobs = _jobs(seeds, res_override=256, spp_pass=2)
    g = Batch(graph_jobs, use_graph=True)
    e = Batch(eager_jobs, use_graph=False)
    passes = g.passes_needed
    assert passes >= 3
    scratch = torch.rand(1 << 22, device="cuda")
    for k in range(passes):
        g.run(1)
        e.run(1)
        total = float((scratch * (k + 1)).sum())
        torch.cuda.synchronize()
        assert np.isfinite(total)
        for gj, ej in zip(graph_jobs, eager_jobs):
            assert np.array_equal(_color(gj), _color(ej)), f"replay {k}: job {gj.seed} differs"
        assert g.rays() == e.rays(), f"replay {k}: ray counts differ"
        print(f"replay {k}: {len(seeds)} images equal, {g.rays()} rays", flush=True)
    kernels, memsets, others = g.graph_nodes()
    print(f"captured pass: {kernels} kernel nodes, {memsets} memset nodes, {others} other nodes", flush=True)
    assert kernels > 0 and memsets == 0 and others == 0
    assert e.graph_nodes() == (-1, -1, -1)
    g.close()
    e.close()
