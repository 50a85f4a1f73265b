"""Benchmark: path-tracing throughput on one MI355X.

A step renders the large configuration once: 500 seeded displaced UV
spheres (2,000,000 triangles), 64 analytic spheres and three sphere lights
in a 100^3 box, 1920x1080, 16 samples per pixel, depth 6, next-event
estimation (--config picks another workload).  Scene flattening and the
GPU BVH build happen once before the timed region; inputs are resident in
HBM.

Prints one JSON line:
  value        Mrays/s over the timed steps (rays counted as the reference
               counts them: one per bounce-loop iteration plus one per
               shadow ray cast)
  roofline     the closest-hit (extend) kernel: HBM bytes per launch
               (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE, committed under
               profiles/) over its average launch time measured with HIP
               events in the timed region, against 8 TB/s; the 64/36/16-byte
               record rate beside it as algorithmic_gbs
  rmse         against the fp64 oracle at 4096 spp on the fixture's pixels
  configs      the other single-scene workloads (Cornell, mesh), the same way
  batching     256 uneven scenes in one batch against one request at a time
  serving      Poisson arrivals: the continuous-batching server against the
               one-request-at-a-time loop, p50/p99 latency
  cpu_baseline the CPU oracle (fp32 restatement of the reference kernel)
               on a bounded pixel subset of the headline workload
"""
from __future__ import annotations

import argparse
import dataclasses
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
NODE_B, TRI_B, PRIM_B = 64, 36, 16


def _load_traffic(name, samples_per_pass):
    """HBM bytes per extend launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel of known size), if recorded for the
    same workload and pass size (a launch's traffic grows with the pass),
    and the counters that say what binds the kernel."""
    p = ROOT / "profiles" / f"traffic_{name}.json"
    if p.exists():
        try:
            t = json.loads(p.read_text())
        except Exception:
            return None, None
        if t.get("samples_per_pass") == samples_per_pass:
            return t.get("bytes_per_launch"), t.get("binding")
    return None, None


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(name, target_seconds=15.0, threads=None):
    """Oracle fp32 render of a pixel subset of the same workload."""
    from oracle import oracle as orc
    from roxy_dcc_amd import renderer as R
    from roxy_dcc_amd import scenes
    from roxy_dcc_amd.extract import flatten_scene
    from tests import helpers

    world, W, H, st = scenes.CONFIGS[name]()
    flat = flatten_scene(world)
    osc = helpers.flat_to_oracle(flat)
    cam = world.active_camera
    ocam = helpers.oracle_camera(cam, W, H)
    mode = R.TaichiRenderer._direct_light_mode_to_id(st["direct_light_mode"])
    prm = orc.params(st["max_depth"], int(world._use_sky), mode, st["direct_light_max_depth"], 10.0,
                     world._background_color, 0)
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count() or 1
    rng = np.random.default_rng(0)
    n_pix = 256
    spp = st["samples"]
    elapsed, rays, pix_done = 0.0, 0, 0
    while elapsed < target_seconds and pix_done < W * H:
        pix = rng.choice(W * H, size=n_pix, replace=False).astype(np.int32)
        acc = orc.Accum(n_pix)
        t0 = time.perf_counter()
        orc.render(osc, ocam, prm, 0, spp, acc, threads=threads, pixels=pix)
        elapsed += time.perf_counter() - t0
        rays += acc.counts.rays
        pix_done += n_pix
        n_pix = min(n_pix * 2, 65536)
    return {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": f"{pix_done} random pixels x {spp} spp of the {W}x{H} {name} frame, fp32 oracle, "
                      f"{elapsed:.1f} s"}


def _roofline(work, prof, steps, traffic, pmc=None):
    """k_extend's roofline: the HBM bytes one launch moves (rocprofv3
    FETCH_SIZE x2 + WRITE_SIZE, from the committed profile of this workload
    at this pass size) over its average launch time measured live with HIP
    events, against MI355X's 8 TB/s.  The 64/36/16-byte count of node,
    triangle and primitive records the traversal consumes (most served by
    L2 and the Infinity Cache) is reported beside it as
    ``algorithmic_gbs``, not as a fraction of HBM.  When the primitive scan
    ran as its own pre-pass (k_prim_scan) its tests are not k_extend's
    work."""
    alg = NODE_B * work["node_pops"] + TRI_B * work["tri_tests"]
    if prof.n_prim_scan == 0:
        alg += PRIM_B * work["prim_tests"]
    launches = prof.n_extend / steps
    alg_per_launch = alg / max(launches, 1)
    ms = prof.ms_extend / max(prof.n_extend, 1)
    achieved = traffic / (ms * 1e-3) / 1e9 if traffic else None
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
           "kernel": "k_extend (closest-hit traversal)", "avg_launch_ms": ms, "launches_per_step": launches,
           "algorithmic_bytes_per_launch": alg_per_launch,
           "algorithmic_gbs": alg_per_launch / (ms * 1e-3) / 1e9}
    if pmc:
        out["binding"] = pmc
    return out


def _stages(prof, steps):
    return {k: getattr(prof, "ms_" + k) / steps for k in
            ("generate", "prim_scan", "reorder", "extend", "shade", "connect", "accumulate")}


def bvh_build_times(ex, reps=5):
    """GPU BVH builds over the scene's resident triangles: the Morton build
    (bvh_builder.py:412-486) and core/bvh.py's object-median rule."""
    import torch

    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd.bvh import GPUBVHBuilder

    b = GPUBVHBuilder()
    tb = ex.flat.triangles()
    b.build([tb])
    torch.cuda.synchronize()
    ws = C.c_size_t()
    n = tb.triangle_count
    nat.check(nat.lib().rx_bvh_workspace_size(n, 8, C.byref(ws)))
    work_buf = torch.empty(ws.value, dtype=torch.uint8, device="cuda")
    ref = b._ref_struct()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        nat.check(nat.lib().rx_bvh_build_morton(b.d["v0"].data_ptr(), b.d["v1"].data_ptr(), b.d["v2"].data_ptr(), n,
                                                8, work_buf.data_ptr(), ws.value, C.byref(ref),
                                                b.d["nodes"].data_ptr(), b.d["header"].data_ptr(),
                                                b.d["order"].data_ptr(), nat.stream_ptr()))
    ev1.record()
    torch.cuda.synchronize()
    build_ms = ev0.elapsed_time(ev1) / reps
    om = GPUBVHBuilder()
    om._reset_for_build(tb)
    om._build_object_median_gpu()
    v64 = [om.d[k].double().contiguous() for k in ("v0", "v1", "v2")]
    om_ws = C.c_size_t()
    nat.check(nat.lib().rx_bvh_object_median_workspace_size(n, om.OBJECT_MEDIAN_LEAF_SIZE, C.byref(om_ws)))
    om_buf = torch.empty(om_ws.value, dtype=torch.uint8, device="cuda")
    om_ref = om._ref_struct()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        nat.check(nat.lib().rx_bvh_build_object_median(
            v64[0].data_ptr(), v64[1].data_ptr(), v64[2].data_ptr(), n, om.OBJECT_MEDIAN_LEAF_SIZE,
            om_buf.data_ptr(), om_ws.value, C.byref(om_ref), om.d["nodes"].data_ptr(), om.d["header"].data_ptr(),
            om.d["order"].data_ptr(), nat.stream_ptr()))
    om_ms = (time.perf_counter() - t0) * 1e3 / reps
    return {"triangles": n, "morton_nodes": b.node_count, "morton_depth": b.depth, "gpu_morton_build_ms": build_ms,
            "gpu_object_median_build_ms": om_ms, "object_median_nodes_depth": (om.node_count, om.depth)}


def _rmse(name, frame):
    """RMSE of the linear accumulator against the fp64 oracle's 4096-spp means
    on the fixture's pixels (tests/golden/rmse_<name>.npz)."""
    p = ROOT / "tests" / "golden" / f"rmse_{name}.npz"
    if not p.exists():
        return None
    ref = np.load(p)
    got = frame.color.cpu().numpy().reshape(-1, 3)[ref["pixels"]].astype(np.float64)
    err = got - ref["color"]
    return {"rmse": float(np.sqrt(np.mean(err ** 2))), "mean_error": float(err.mean()),
            "pixels": int(len(ref["pixels"])), "reference_spp": int(ref["spp"])}


def config_measure(name, steps, warmup=1, extra_flags=0):
    """One single-scene workload: `steps` timed renders (stage times from HIP
    events), the traversal work of one counted render, RMSE."""
    import torch

    from roxy_dcc_amd import _native as nat
    from roxy_dcc_amd import renderer as R
    from roxy_dcc_amd import scenes
    from roxy_dcc_amd.denoise import edge_aware_denoise, linear_to_display
    from roxy_dcc_amd.extract import extract_scene

    world, W, H, st = scenes.CONFIGS[name]()
    t0 = time.perf_counter()
    ex = extract_scene(world)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    sc = ex.scene
    spp = st["samples"]
    per_pass = R.default_samples_per_pass(W, H, spp)
    integ = R.Integrator(sc, W, H, per_pass)
    frame = R.DeviceFrame(W, H)
    cam = R.camera_struct(world.active_camera, W, H)
    mode = R.TaichiRenderer._direct_light_mode_to_id(st["direct_light_mode"])
    denoise_passes = st.get("denoise_passes", 0)

    def step(flags):
        prm = R.params_struct(st["max_depth"], int(world._use_sky), mode, st["direct_light_max_depth"], 10.0,
                              world._background_color, 0, per_pass, flags | extra_flags)
        done = 0
        while done < spp:
            n = min(per_pass, spp - done)
            integ.render(frame, cam, prm, done, n)
            done += n
        if denoise_passes:
            x = frame.color
            for k in range(denoise_passes):
                x = edge_aware_denoise(x, 1, 0.08, 0.8, normal=frame.normal, albedo=frame.albedo,
                                       depth=frame.depth, dilation=1 << k)
            linear_to_display(x)

    for _ in range(warmup):
        step(0)
    integ.reset_counts()
    step(nat.RX_FLAG_COUNT_TRAFFIC)
    torch.cuda.synchronize()
    work = integ.counts()
    integ.reset_counts()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(nat.RX_FLAG_PROFILE)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    counts = integ.counts()
    traffic, pmc = _load_traffic(name, per_pass)
    return {
        "workload": f"{name}: {W}x{H}, {spp} spp, depth {st['max_depth']}, NEE {st['direct_light_mode']}, "
                    f"{sc.n_tris} triangles, {sc.n_prims} primitives, {sc.n_lights} lights, "
                    f"{denoise_passes} a-trous denoise passes",
        "steps": steps, "dt": dt, "rays": counts["rays"],
        "Mrays_per_s": counts["rays"] / dt / 1e6, "ms_per_render": dt * 1e3 / steps,
        "renders_per_s": steps / dt, "samples_per_s": W * H * spp * steps / dt,
        "rays_per_render": counts["rays"] // steps, "samples_per_pass": per_pass,
        "stage_ms_per_render": _stages(integ.profile, steps), "scene_setup_s": setup,
        "roofline": _roofline(work, integ.profile, steps, traffic, pmc),
        "traversal_work_per_render": work, "rmse": _rmse(name, frame),
        "_ex": ex,
    }


def _plain_render(R, torch, world, W, H, st, sc, seed, per_sample=False):
    """One request through the single-job path, then the frame read back.
    per_sample: the reference renderer's execution model, one launch
    sequence per sample (rendering/taichi_renderer.py:140-156); otherwise
    the drop-in's policy of passes of up to 2^25 paths."""
    spp = 1 if per_sample else R.default_samples_per_pass(W, H, st["samples"])
    integ = R.Integrator(sc, W, H, spp)
    frame = R.DeviceFrame(W, H)
    cam = R.camera_struct(world.active_camera, W, H)
    mode = R.TaichiRenderer._direct_light_mode_to_id(st["direct_light_mode"])
    prm = R.params_struct(st["max_depth"], int(world._use_sky), mode, st["direct_light_max_depth"], 10.0,
                          world._background_color, seed, spp)
    done = 0
    while done < st["samples"]:
        n = min(spp, st["samples"] - done)
        integ.render(frame, cam, prm, done, n)
        done += n
    return frame.color.cpu().numpy(), integ.counts()["rays"]


def batching_bench(n_scenes=256, seed0=4000):
    """Many uneven scenes (displaced spheres of 1k..200k triangles, 64^2..512^2,
    4..32 spp, depth 6) rendered one request at a time through the plain path,
    then as one autotuned batch (one set of launches per pass for all jobs,
    replayed from a HIP graph).  Scenes are resident on both sides; both
    include reading every image back.  Images must be bit-identical, so
    the two are compared at equal RMSE by construction."""
    import torch

    from roxy_dcc_amd import renderer as R
    from roxy_dcc_amd import scenes
    from roxy_dcc_amd.autotune import Autotuner
    from roxy_dcc_amd.batch import Batch, RenderJob
    from roxy_dcc_amd.extract import extract_scene

    t0 = time.perf_counter()
    items = []
    for i in range(n_scenes):
        world, W, H, st = scenes.batch_scene(seed0 + i)
        items.append((world, W, H, st, extract_scene(world).scene))
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    tris = sum(it[4].n_tris for it in items)
    pixels = sum(it[1] * it[2] for it in items)
    samples = sum(it[1] * it[2] * it[3]["samples"] for it in items)

    _plain_render(R, torch, *items[0][:5], seed0)          # warm-up
    plain, plain_s, plain_rays = {}, {}, {}
    for mode in ("per_sample", "drop_in"):
        t0 = time.perf_counter()
        imgs, rays = [], 0
        for i, (world, W, H, st, sc) in enumerate(items):
            img, r = _plain_render(R, torch, world, W, H, st, sc, seed0 + i, per_sample=mode == "per_sample")
            imgs.append(img)
            rays += r
        plain_s[mode] = time.perf_counter() - t0
        plain[mode], plain_rays[mode] = imgs, rays

    jobs = [RenderJob(sc, world.active_camera, W, H, st["samples"], max_depth=st["max_depth"],
                      direct_light_mode=st["direct_light_mode"], direct_light_max_depth=st["direct_light_max_depth"],
                      use_sky=bool(world._use_sky), background=tuple(world._background_color), seed=seed0 + i)
            for i, (world, W, H, st, sc) in enumerate(items)]
    tuner = Autotuner()
    t0 = time.perf_counter()
    tuning = tuner.tune(jobs)
    tune_s = time.perf_counter() - t0
    tuning.apply(jobs)
    t0 = time.perf_counter()
    b = Batch(jobs, use_graph=True)
    b.run()
    torch.cuda.synchronize()
    batched = [j.frame.color.cpu().numpy() for j in jobs]
    batch_s = time.perf_counter() - t0
    rays = b.rays()
    b.close()
    same = {m: sum(int(np.array_equal(a, c)) for a, c in zip(plain[m], batched)) for m in plain}
    return {"workload": f"{n_scenes} scenes (seeds {seed0}..{seed0 + n_scenes - 1}): displaced UV spheres of 1k..200k "
                        "triangles in the Cornell enclosure, 64^2..512^2, 4..32 spp, depth 6",
            "triangles": tris, "pixels": pixels, "samples": samples, "scene_setup_s": setup,
            "batched_s": batch_s, "batched_renders_per_s": n_scenes / batch_s,
            "batched_Mrays_per_s": rays / batch_s / 1e6,
            "plain_per_sample_s": plain_s["per_sample"],
            "plain_per_sample_renders_per_s": n_scenes / plain_s["per_sample"],
            "speedup_vs_plain_per_sample": plain_s["per_sample"] / batch_s,
            "plain_drop_in_s": plain_s["drop_in"], "plain_drop_in_renders_per_s": n_scenes / plain_s["drop_in"],
            "plain_drop_in_Mrays_per_s": plain_rays["drop_in"] / plain_s["drop_in"] / 1e6,
            "speedup_vs_plain_drop_in": plain_s["drop_in"] / batch_s,
            "identical_images": same, "tuning": dataclasses.asdict(tuning), "tune_s": tune_s,
            "tune_trials": [[dataclasses.asdict(t), r / 1e6] for t, r in tuner.log]}


def serving_bench(n_requests=40, loads=(0.5, 1.0, 2.0, 4.0), seed=5):
    """Poisson arrivals of render requests drawn from the mesh (70k
    triangles) and large (2M triangles) scene generators, 128^2..1080p,
    4..64 spp, served by the continuous-batching server and by the
    one-request-at-a-time loop.  Offered load is a multiple of the serial
    loop's capacity (requests/s it sustains with every request queued at
    t = 0).  Each request's image must be the same in both."""
    import torch

    from roxy_dcc_amd import scenes
    from roxy_dcc_amd.extract import extract_scene
    from roxy_dcc_amd.serve import RenderServer, Request, SceneCache, latency_stats, poisson_arrivals, serial_loop

    def load(key):
        kind, sd = key
        world, _, _, st = scenes.mesh_scene(seed=sd) if kind == "mesh" else scenes.large_scene(seed=sd)
        return extract_scene(world).scene, world.active_camera, {
            "use_sky": bool(world._use_sky), "background": tuple(world._background_color),
            "direct_light_mode": st["direct_light_mode"], "direct_light_max_depth": st["direct_light_max_depth"]}

    rng = np.random.default_rng(seed)
    keys = [("mesh", 20 + k) for k in range(6)] + [("large", 30 + k) for k in range(2)]
    sizes = [(128, 128), (256, 256), (512, 512), (960, 540), (1920, 1080)]
    spec = []
    for i in range(n_requests):
        key = keys[int(rng.integers(len(keys)))]
        W, H = sizes[int(rng.choice(len(sizes), p=[0.3, 0.3, 0.2, 0.15, 0.05]))]
        # a quarter of the requests carry an error target (the adaptive stop
        # of rendering/ray_tracer.py:531-541, threshold 0.005, at most 64 spp)
        if rng.random() < 0.25:
            spec.append((key, W, H, 64, (0.005, 4, 1)))
        else:
            spec.append((key, W, H, int(rng.choice([4, 8, 16, 32, 64])), None))

    def requests(arrivals):
        return [Request(i, key, W, H, spp, float(a), max_depth=6, seed=1000 + i, adaptive=ad)
                for i, ((key, W, H, spp, ad), a) in enumerate(zip(spec, arrivals))]

    cache = SceneCache(load, max_scenes=32)
    t0 = time.perf_counter()
    for k in keys:
        cache.get(k)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    serial_loop(cache, requests([0.0] * 2)[:2])               # warm-up
    t0 = time.perf_counter()
    ref = serial_loop(cache, requests([0.0] * n_requests))
    cap = n_requests / (time.perf_counter() - t0)
    ref_frames = {r.rid: r.frame for r in ref}
    adaptive_spp = [r.samples_done for r in ref if spec[r.rid][4] is not None]
    rows = []
    for f in loads:
        rate = f * cap
        arr = poisson_arrivals(rate, n_requests, seed=seed)
        row = {"offered_load": f, "rate_per_s": rate}
        t0 = time.perf_counter()
        done = serial_loop(cache, requests(arr))
        row["serial"] = latency_stats(done, time.perf_counter() - t0)
        t0 = time.perf_counter()
        done = serial_loop(cache, requests(arr), paths_per_pass=1)
        row["serial_per_sample"] = latency_stats(done, time.perf_counter() - t0)
        same = 0
        for policy in ("deadline", "srpt", "fifo"):
            server = RenderServer(cache, max_jobs=32, policy=policy)
            t0 = time.perf_counter()
            done = server.run(requests(arr))
            row["server_" + policy] = dict(latency_stats(done, time.perf_counter() - t0), passes=server.passes,
                                           pass_s=server.pass_s)
            same += sum(int(np.array_equal(r.frame, ref_frames[r.rid])) for r in done)
        row["identical_images"] = same
        rows.append(row)
    return {"workload": f"{n_requests} requests over {len(keys)} resident scenes (6 mesh, 2 large), "
                        "128^2..1920x1080, 4..64 spp or an error target (adaptive stop, threshold 0.005), depth 6, "
                        "Poisson arrivals; offered load relative to the "
                        "serial loop's capacity; serial = the drop-in renderer (passes of up to 2^25 paths), serial_per_sample = one "
                        "sample per pass (the reference's launch-per-sample model); identical_images counts "
                        "the three server policies against the serial images",
            "scene_setup_s": setup, "serial_capacity_per_s": cap, "error_target_requests": len(adaptive_spp),
            "error_target_samples_taken": adaptive_spp, "sweep": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="large", choices=["cornell", "mesh", "large"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-configs", action="store_true", help="skip the other single-scene workloads")
    ap.add_argument("--no-bvh", action="store_true", help="skip the BVH build timings")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-scene batching comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (Poisson arrivals) comparison")
    ap.add_argument("--prim-prepass", action="store_true",
                    help="run the closest-primitive scan as its own pass (RX_FLAG_PRIM_PREPASS)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(pixels/samples would shard across GPUs with an accumulator all-reduce)", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        n = args.copy_probe << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "copies": 5}))
        return

    from roxy_dcc_amd import _native as nat

    nat.require_gpu()
    extra_flags = nat.RX_FLAG_PRIM_PREPASS if args.prim_prepass else 0
    head = config_measure(args.config, args.steps, args.warmup, extra_flags)
    ex = head.pop("_ex")
    out = {
        "metric": "Mrays/s",
        "value": head["Mrays_per_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_render"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": head["workload"], "samples_per_pass": head["samples_per_pass"],
                   "parallelism": "single GPU"},
        "renders_per_s": head["renders_per_s"],
        "samples_per_s": head["samples_per_s"],
        "rays_per_step": head["rays_per_render"],
        "stage_ms_per_step": head["stage_ms_per_render"],
        "scene_setup_s": head["scene_setup_s"],
        "roofline": head["roofline"],
        "rmse": head["rmse"],
        "traversal_work_per_step": head["traversal_work_per_render"],
    }
    if not args.no_bvh:
        out["bvh_build"] = bvh_build_times(ex)
    del ex
    if not args.no_configs:
        extra = {}
        for name in ("cornell", "mesh", "large"):
            if name != args.config:
                r = config_measure(name, 3 if name != "cornell" else 10, 1, extra_flags)
                r.pop("_ex")
                extra[name] = r
        out["configs"] = extra
    if not args.no_batch:
        out["batching"] = batching_bench()
    if not args.no_serve:
        out["serving"] = serving_bench()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.config, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    # release every device buffer before the interpreter's own teardown
    import gc

    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
