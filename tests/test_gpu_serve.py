"""Continuous batching: every request's image is bit-identical whether it
was rendered alone or in the server's changing mix of jobs."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _loader():
    from roxy_dcc_amd import scenes
    from roxy_dcc_amd.extract import extract_scene

    def load(key):
        world, *_ = scenes.batch_scene(key)
        return extract_scene(world).scene, world.active_camera

    return load


def _requests():
    from roxy_dcc_amd.serve import Request

    spec = [(4100, 32, 5), (4101, 48, 3), (4100, 24, 7), (4102, 40, 4), (4103, 16, 9), (4101, 32, 2)]
    return [Request(i, key, res, res, spp, arrival=0.002 * i, max_depth=6, seed=77 + i)
            for i, (key, res, spp) in enumerate(spec)]


def test_served_results_equal_alone_results():
    from roxy_dcc_amd.serve import RenderServer, SceneCache, latency_stats, serial_loop

    server = RenderServer(SceneCache(_loader(), max_scenes=3), max_jobs=4, pass_paths=4096)
    done = server.run(_requests())
    assert len(done) == 6 and server.passes > 1
    alone = {r.rid: r for r in serial_loop(SceneCache(_loader()), _requests(), paths_per_pass=1 << 20)}
    for r in done:
        same = np.array_equal(r.frame, alone[r.rid].frame)
        print(f"request {r.rid}: in-mix == alone {same}, rays {r.rays} / {alone[r.rid].rays}", flush=True)
        assert same
        assert r.rays == alone[r.rid].rays
    stats = latency_stats(done, max(r.finished for r in done))
    assert stats["completed"] == 6 and stats["p99_ms"] >= stats["p50_ms"] > 0
