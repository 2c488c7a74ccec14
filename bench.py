"""Wrap throughput on one MI355X, beside the FP64 CPU oracle.

A step is one complete wrap of the benchmark job -- device template graph,
target hierarchy, and the whole coarse-to-fine Gauss-Newton schedule -- with
the job's inputs already resident in HBM (the session is created and its
meshes uploaded before the timed region; each step rewinds and re-runs it).

Default workload (--config 2): the 200k-vertex job -- a class-I geodesic sphere
of frequency 141 (198,812 vertices / 397,620 triangles) onto a frequency-200
sphere (800,000 triangles) with 8 seeded RBF bumps and normal noise of 0.05 %
of the box diagonal; 8,000 graph nodes over 3 levels, 8 graph neighbours, 10
Gauss-Newton steps per level. --config 1 is the 50k capsule humanoid (2,000
nodes, 4 levels), --config 0 the 10k icosphere (500 nodes, 3 levels), and
--config 3 the 200k-vertex job built with the humanoid generator instead of
spheres (203k-vertex template, 794k-triangle posed target with holes and the
same normal noise, 8,000 nodes over 3 levels).

Three device forms are timed through the same device schedule: "float"
(the configuration's precision recipe and the headline `value`: assembly
blocks computed and accumulated in FP32, FP32 PCG iterations, FP64
refinement of that FP32 system; right-hand side, energies, decisions and
rotation fits in FP64), "mixed" (FP64 assembly, FP32 PCG iterations refined
against the FP64 system) and "double" (plain FP64). Every system is solved to
a 1e-8 relative true residual (the schedule's PCG tolerance) with the
schedule's cap of 50 PCG iterations; the line reports the most iterations any
solve took. The oracle solves each system directly.

CPU baseline: the FP64 oracle (a single-threaded restatement of the reference's
route A) on this host, run before the GPU is touched. For configs 0 and 1 it
runs the whole wrap, once on one core and then as independent copies on the
usable cores (at most 16, this host's share of the machine). For configs 2
and 3 a whole wrap takes minutes, so the all-core throughput
(`cpu_baseline.value`) comes from a bounded sample -- the same job with one
outer iteration per level, wrap time extrapolated as setup + 10 x the
sample's solve time (the per-iteration cost is flat within a level) -- and
one complete wrap runs on one core in a background process while the GPU
forms are timed: its measured time is reported beside the extrapolation, and
every form's vertex error and final correspondence counts are taken against
its result.

The timed wraps run with profiling off. Two more wraps per form then run with
HIP-event stage timing and traversal counters on; the stage breakdown, the
roofline figures and the query statistics come from those.

Beside the headline (unless --no-extras), in the headline's form: the 64-job
batch of uneven humanoid jobs (nricp_amd.batch: jobs/s as packed lockstep
batches at the autotuner's shape and as concurrent sessions at a fixed slot
count, the same jobs one at a time, and whether a sample of jobs gives the
same bits alone and in the batch), and the serving loop (nricp_amd.serve_bench: p50 /
p95 / p99 latency of Poisson arrivals at two offered loads below saturation).
Their jobs are generated before the GPU is touched; they run while the
one-core oracle wrap is still going.

Prints one JSON line (rank 0).
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for sub in ("", "nricp_amd", "oracle"):
    p = str(ROOT / sub) if sub else str(ROOT)
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
TOLERANCE = 1e-8  # the benchmark schedule's PCG tolerance (--pcg-tolerance)
MAX_PCG = 50       # the configurations' PCG iteration cap (--pcg-max-iterations)
OUTER = 10
# HBM bytes per pcg_kernel launch (FETCH_SIZE + WRITE_SIZE per dispatch, each
# scaled by what the counter reads for a plain device copy of known size),
# averaged over a wrap's launches: measured on the 200k-vertex job by the
# counter passes recorded in profiles/pcg_kernel_traffic_geodesic-200k.md, per
# (config, form, PCG tolerance). Other jobs and tolerances report null.
PCG_TRAFFIC = {(2, "float", 1e-8): 3.185e7}
# Device-timed stages of a wrap (HIP events around each stage's launches).
STAGES = ("fps", "graph", "target_bvh", "layout", "deform", "closest_points", "assemble", "pcg", "energy",
          "local_step")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_job(config: int, seed: int):
    from nricp_amd import workloads

    job = workloads.config_job(config, seed)
    settings = workloads.level_settings(job.graph_levels, outer=OUTER, inner=1)
    kwargs = dict(graph_levels=job.graph_levels, nodes_initial=job.nodes_initial, growth=2, graph_neighbours=8)
    return job, settings, kwargs


def run_form(dp, job, settings, kwargs, form: str, steps: int, warmup: int, profiled: int = 2) -> dict:
    # Both forms run the same device schedule (route B's session, route A
    # semantics) to the same PCG tolerance; only the arithmetic differs.
    session = dp.RouteBSession(job.template_points, job.template_faces, job.target_points, job.target_faces,
                               settings, route="route_b", precision=form,
                               pcg_tolerance=TOLERANCE, pcg_max_iterations=MAX_PCG, **kwargs)
    for w in range(warmup):
        session.reset()
        session.run()
        log(f"{form}: warmup {w + 1}/{warmup} done")
    # Timed wraps with profiling off: the headline numbers.
    before = session.counters()
    dp.device_synchronize()
    latencies = []
    t_start = time.perf_counter()
    for k in range(steps):
        t0 = time.perf_counter()
        session.reset()
        session.run()
        latencies.append(time.perf_counter() - t0)
        log(f"{form}: step {k + 1}/{steps} {latencies[-1] * 1e3:.1f} ms")
    dp.device_synchronize()
    elapsed = time.perf_counter() - t_start
    after = session.counters()
    delta = {k: after[k] - before.get(k, 0.0) for k in after}
    # Then profiled wraps (HIP events around each stage, traversal counters)
    # for the stage breakdown and the roofline figures.
    session.profile(True)
    p_before = session.counters()
    stage_before = {s: session.kernel_time(s) for s in STAGES}
    for k in range(profiled):
        session.reset()
        session.run()
    p_after = session.counters()
    stages = {}
    for s, (ms0, n0) in stage_before.items():
        ms1, n1 = session.kernel_time(s)
        stages[s] = {"ms": (ms1 - ms0) / profiled, "launches": (n1 - n0) / profiled}
    p_delta = {k: (p_after[k] - p_before.get(k, 0.0)) / profiled for k in p_after}
    log(f"{form}: {profiled} profiled wraps done")
    result = session.result()
    report = json.loads(result["report"])
    del session
    return {
        "form": form,
        "solves_per_s": steps / elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "p50_ms": 1e3 * statistics.median(latencies),
        "p95_ms": 1e3 * float(np.percentile(latencies, 95)),
        "p99_ms": 1e3 * float(np.percentile(latencies, 99)),
        "gn_solves_per_s": delta["gauss_newton_solves"] / elapsed,
        "counters": delta,
        "profiled_counters": p_delta,
        "stages": stages,
        "vertices": np.asarray(result["vertices"]),
        "node_positions": np.asarray(result["node_positions"]),
        "level_nodes": [int(n) for n in result["level_nodes"]],
        "report": report,
        "max_query_steps": p_after.get("query_max_steps", 0.0),
        "profiled_counters_max": after.get("pcg_max_iterations", 0.0),
    }


def hbm_roofline(form: dict, stage: str, bytes_key: str, kernel: str, traffic=None) -> dict:
    """Algorithmic bytes of one launch over its average device time, both
    from the profiled wraps (per wrap)."""
    st = form["stages"][stage]
    seconds = st["ms"] * 1e-3
    launches = max(st["launches"], 1)
    bytes_per_launch = form["profiled_counters"][bytes_key] / launches
    achieved = (bytes_per_launch / (seconds / launches)) / 1e9 if seconds > 0 else 0.0
    out = {
        "kernel": kernel,
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "avg_launch_us": round(1e6 * seconds / launches, 2),
        "algorithmic_bytes_per_launch": round(bytes_per_launch),
    }
    if stage == "pcg":
        out["iterations_per_launch"] = round(form["profiled_counters"]["pcg_iterations"] / launches, 2)
    return out


def query_stats(form: dict) -> dict:
    c = form["profiled_counters"]
    rows = max(c.get("query_rows", 0.0), 1.0)
    ms = form["stages"]["closest_points"]["ms"]
    return {
        "queries_per_s": round(rows / (ms * 1e-3), 1) if ms > 0 else None,
        "node_opens_per_query": round(c.get("query_node_opens", 0.0) / rows, 2),
        "triangle_tests_per_query": round(c.get("query_leaf_tests", 0.0) / rows, 2),
        "max_node_opens_one_query": int(form["max_query_steps"]),
    }


def _oracle_full_wrap(job, settings, kwargs) -> dict:
    """One complete oracle wrap (run in a worker process)."""
    import nricp_oracle

    t0 = time.perf_counter()
    result = nricp_oracle.solve_route_a(job.template_points, job.template_faces, job.target_points, job.target_faces,
                                        settings, progress=True, **kwargs)
    report = json.loads(result["report"])
    return {"wall": time.perf_counter() - t0, "vertices": np.asarray(result["vertices"]),
            "correspondence": report.get("correspondence")}


def _oracle_sample(config: int, job, settings, kwargs) -> dict:
    """One oracle run of the CPU baseline's sample (in this process)."""
    import nricp_oracle

    full = config in (0, 1)
    sample_settings = settings if full else [dict(l, max_outer_iterations=1) for l in settings]
    t0 = time.perf_counter()
    result = nricp_oracle.solve_route_a(job.template_points, job.template_faces, job.target_points, job.target_faces,
                                        sample_settings, pcg_tolerance=TOLERANCE, pcg_max_iterations=MAX_PCG,
                                        **kwargs)
    wall = time.perf_counter() - t0
    report = json.loads(result["report"])
    seconds = wall if full else report["setup_seconds"] + OUTER * report["solve_seconds"]
    # The template graph is built in full either way, so the finest level's
    # nodes compare at every size; the final correspondences only after a
    # complete wrap.
    return {"wall": wall, "seconds": seconds, "setup": report["setup_seconds"], "solve": report["solve_seconds"],
            "vertices": np.asarray(result["vertices"]) if full else None,
            "node_positions": np.asarray(result["node_positions"]),
            "level_nodes": [int(n) for n in result["level_nodes"]],
            "correspondence": report.get("correspondence") if full else None}


def _host_description() -> str:
    model = "unknown CPU"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        smt = "SMT on" if Path("/sys/devices/system/cpu/smt/active").read_text().strip() == "1" else "SMT off"
    except OSError:
        smt = "SMT unknown"
    return f"{model}, {smt}"


def cpu_baseline(config: int, job, settings, kwargs, procs: int) -> dict:
    """The FP64 oracle on this host: one job on one core (the accuracy
    reference and the latency), then `procs` independent copies at once, one
    per core, for the all-core throughput. Runs before the GPU is touched, so
    the worker processes start from a process with no device state."""
    full = config in (0, 1)
    one = _oracle_sample(config, job, settings, kwargs)
    what = (f"one complete wrap of {job.name} ({job.graph_levels} levels x {OUTER} outer iterations)" if full else
            f"{job.name} with 1 outer iteration per level (of {OUTER}), wrap time extrapolated as setup + {OUTER} x "
            f"solve")
    log(f"cpu oracle, 1 core: {one['wall']:.1f} s measured, {one['seconds']:.1f} s per wrap")
    per_core = [one["seconds"]]
    if procs > 1:
        import multiprocessing as mp

        ctx = mp.get_context("spawn")
        t0 = time.perf_counter()
        with ctx.Pool(procs) as pool:
            runs = pool.starmap(_oracle_sample, [(config, job, settings, kwargs)] * procs)
        log(f"cpu oracle, {procs} processes at once: {time.perf_counter() - t0:.1f} s wall")
        per_core = [r["seconds"] for r in runs]
    throughput = float(sum(1.0 / t for t in per_core))
    sample = (f"FP64 route-A restatement (single-threaded per job) on {_host_description()}: {what}; 1 core "
              f"{one['seconds']:.1f} s per wrap (measured {one['wall']:.1f} s); {len(per_core)} independent jobs on "
              f"{len(per_core)} cores at once: {throughput:.4f} wraps/s")
    return {
        "value": throughput,
        "unit": "solves/s",
        "cores": len(per_core),
        "kind": "port",
        "sample": sample,
        "one_core_solves_per_s": 1.0 / one["seconds"],
        "seconds_per_wrap_one_core": one["seconds"],
        "vertices": one["vertices"],
        "node_positions": one["node_positions"],
        "level_nodes": one["level_nodes"],
        "correspondence": one["correspondence"],
    }


CPU_ARRAYS = ("vertices", "node_positions", "level_nodes", "correspondence")


BATCH_JOBS, BATCH_SEED = 64, 1000        # nricp_amd.batch's defaults
PACKED_SHAPE = (2, 16, 4)                 # PCG node slots per lane group, jobs per packed batch, batches at once
SERVING_POOL, SERVING_SEED = 16, 3000     # nricp_amd.serve_bench's defaults


def generate_extras(big) -> dict:
    """The batch's 64 jobs and the serving pool, generated in worker
    processes before this process touches the GPU."""
    from nricp_amd import batch as batch_mod
    from nricp_amd import serve_bench

    t0 = time.perf_counter()
    jobs = batch_mod.make_batch(BATCH_JOBS, BATCH_SEED)
    pool = serve_bench.build_pool(SERVING_POOL, SERVING_SEED, min(len(os.sched_getaffinity(0)), 16))
    log(f"batch ({len(jobs)} jobs) and serving pool ({len(pool)} jobs) generated in "
        f"{time.perf_counter() - t0:.1f} s")
    return {"batch": jobs, "pool": pool, "big": big}


def run_extras(dp, form: str, gen: dict, slots: int, rates: str) -> dict:
    """The 64-job batch and the serving loop, in the headline's form."""
    from nricp_amd import batch as batch_mod
    from nricp_amd import serve_bench

    jobs = gen["batch"]
    batch_mod.run_batch(dp, jobs[:8], 4, form)  # warm-up: code objects, allocator cache
    batch_mod.run_packed(dp, jobs[:8], form, 2, 8)
    # Packed lockstep batches at the shape nricp_amd.batch's autotuner picks
    # on this workload (128-node PCG tile, 4 batches of 16 jobs at once).
    pslots, pchunk, pconc = PACKED_SHAPE
    packed_wall, packed, engine = batch_mod.run_packed(dp, jobs, form, pslots, pchunk, pconc)
    check = min(4, len(jobs))
    packed_alone = [batch_mod.run_packed(dp, [jobs[k]], form, pslots, 1)[1][0] for k in range(check)]
    packed_same = all(np.array_equal(packed[k]["vertices"], packed_alone[k]["vertices"]) for k in range(check))
    wall, together = batch_mod.run_batch(dp, jobs, slots, form)
    serial_wall, alone = batch_mod.run_batch(dp, jobs, 1, form)
    same = all(np.array_equal(together[k]["vertices"], alone[k]["vertices"]) for k in range(check))
    sizes = [len(j.template_points) for j in jobs]
    latency = np.array([r["queued_seconds"] + r["run_seconds"] for r in together]) * 1e3
    log(f"batch: packed {len(jobs) / packed_wall:.1f} jobs/s, {len(jobs) / wall:.1f} jobs/s as concurrent sessions "
        f"at {slots} slots, {len(jobs) / serial_wall:.1f} jobs/s one at a time")
    batch = {
        "metric": "batched uneven instances: jobs/s on one MI355X (whole batch)",
        "workload": f"{len(jobs)} humanoid jobs, {min(sizes)}-{max(sizes)} template vertices, target 2-3x the "
                    f"template's triangles, nodes ~ V/25, 4 levels x {OUTER} GN steps, seed {BATCH_SEED}",
        "form": form,
        "jobs_per_s": round(len(jobs) / packed_wall, 3),
        "packed": {"pcg_slots": pslots, "jobs_per_batch": pchunk, "batches_at_once": pconc,
                   "jobs_per_s": round(len(jobs) / packed_wall, 3),
                   "ticks": int(engine.get("ticks", 0)), "graph_replays": int(engine.get("graph_replays", 0)),
                   "same_bits_alone_and_in_batch": {"jobs_checked": check, "same": bool(packed_same)}},
        "sessions": {"slots": slots, "jobs_per_s": round(len(jobs) / wall, 3),
                     "latency_p50_ms": round(float(np.percentile(latency, 50)), 2),
                     "latency_p99_ms": round(float(np.percentile(latency, 99)), 2),
                     "same_bits_alone_and_in_batch": {"jobs_checked": check, "same": bool(same)}},
        "serial_jobs_per_s": round(len(jobs) / serial_wall, 3),
        "packed_over_sessions": round(wall / packed_wall, 3),
        "packed_over_serial": round(serial_wall / packed_wall, 3),
    }
    pool, big = gen["pool"], gen["big"]
    loads = []
    with dp.Server(max_active=4) as server:
        for j in pool + [big]:  # warm-up: every job once
            a, kw = serve_bench.job_args(j)
            server.result(server.submit(*a, precision=form, **kw))
        for k, rate in enumerate(float(r) for r in rates.split(",") if r.strip()):
            rng = np.random.default_rng([SERVING_SEED, k])
            loads.append(serve_bench.run_load(server, pool, big, rate, 40, form, rng))
            log(f"serving at {rate:g} jobs/s: p99 {loads[-1]['p99_ms']} ms")
    serving = {
        "metric": "serving: p50/p95/p99 latency (submission to result) vs offered load, 1xMI355X",
        "workload": f"Poisson arrivals, 40 per load; {len(pool)} uneven humanoid jobs (2k-50k template vertices) + "
                    f"5 % the 200k-vertex job; 4 jobs on the device at once",
        "loads": loads,
    }
    return {"batch": batch, "serving": serving}


def copy_calibration(mib: int) -> None:
    """Three device-to-device copies of `mib` MiB through the HIP runtime the
    library uses (its blit copy kernel): a plain copy of known size, against
    which rocprofv3's FETCH_SIZE / WRITE_SIZE counts are scaled."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    size = ctypes.c_size_t(mib << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), size) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, size) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(3):
            if hip.hipMemcpy(dst, src, size, 3) != 0:  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    log(f"copied {mib} MiB device to device 3 times (counter calibration)")


def main(argv=None) -> int:
    global TOLERANCE, MAX_PCG
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2,
                    help="0: icosphere 10k, 1: humanoid 50k, 2: geodesic 200k, 3: humanoid 200k")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--forms", default="float,mixed,double")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-oracle-wrap", action="store_true",
                    help="configs 2 and 3: skip the complete one-core oracle wrap (the accuracy reference)")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="oracle jobs run at once for the all-core CPU throughput (default: the usable cores, at "
                         "most 16)")
    ap.add_argument("--pcg-tolerance", type=float, default=TOLERANCE,
                    help="relative true-residual tolerance of every device solve (default 1e-8, the benchmark "
                         "schedule's; 1e-12 keeps both forms on the FP64 oracle's accept/reject trajectory exactly)")
    ap.add_argument("--pcg-max-iterations", type=int, default=MAX_PCG,
                    help="PCG iteration cap per solve (default 50, the benchmark schedule's)")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--no-extras", action="store_true", help="skip the 64-job batch and the serving sample")
    ap.add_argument("--batch-slots", type=int, default=12, help="jobs on the device at once for the batch sample")
    ap.add_argument("--serving-rates", default="20,50", help="offered loads (jobs/s) of the serving sample")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first copy this many MiB device to device three times: a plain copy of known size, against "
                         "which rocprofv3 FETCH_SIZE / WRITE_SIZE counts are scaled")
    args = ap.parse_args(argv)
    TOLERANCE = args.pcg_tolerance
    MAX_PCG = args.pcg_max_iterations

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (independent wrap jobs shard across "
              "GPUs with no collectives; this build runs and measures the one-GPU share)", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))
    extras = not args.no_extras and rank == 0
    if extras:
        # Concurrent jobs' streams each need a hardware queue of their own
        # (HIP reads this once, when it initialises; the default is 4).
        os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

    t0 = time.perf_counter()
    job, settings, kwargs = build_job(args.config, args.seed)
    log(f"job {job.name}: template {len(job.template_points)} v / {len(job.template_faces)} f, target "
        f"{len(job.target_points)} v / {len(job.target_faces)} f, generated in {time.perf_counter() - t0:.1f} s")

    # The CPU baseline first: its worker processes start before this process
    # touches the GPU.
    cpu = None
    full_wrap = full_pool = None
    if not args.no_cpu and rank == 0:
        # At most 16 jobs: this host's share of the machine for one GPU (the
        # affinity mask can show more CPUs than the share).
        procs = args.cpu_procs if args.cpu_procs > 0 else min(len(os.sched_getaffinity(0)), 16)
        cpu = cpu_baseline(args.config, job, settings, kwargs, procs)
        if args.config not in (0, 1) and not args.no_oracle_wrap:
            # The complete wrap runs on one core beside the GPU forms; its
            # process starts before this one touches the GPU.
            import multiprocessing as mp

            full_pool = mp.get_context("spawn").Pool(1)
            full_wrap = full_pool.apply_async(_oracle_full_wrap, (job, settings, kwargs))

    extra_jobs = generate_extras(job) if extras else None

    import nricp_amd

    dp = nricp_amd.load()
    info = dp.device_info()
    if not info["available"]:
        print(f"no HIP device: {info['device']}", file=sys.stderr)
        return 1
    log(f"device: {info['device']}")
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)

    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    results = {f: run_form(dp, job, settings, kwargs, f, args.steps, args.warmup) for f in forms}
    extra_out = run_extras(dp, forms[0], extra_jobs, args.batch_slots, args.serving_rates) if extras else None
    if full_wrap is not None:
        log("waiting for the complete oracle wrap (one core)")
        done = full_wrap.get()
        full_pool.close()
        full_pool.join()
        log(f"complete oracle wrap: {done['wall']:.1f} s on one core")
        cpu["vertices"] = done["vertices"]
        cpu["correspondence"] = done["correspondence"]
        cpu["measured_seconds_per_wrap_one_core"] = done["wall"]
        cpu["sample"] += f"; one complete wrap measured on one core in the same run: {done['wall']:.1f} s"
    head = results[forms[0]]
    diag = float(np.linalg.norm(job.template_points.max(0) - job.template_points.min(0)))

    def error(a, b):
        e = np.linalg.norm(a - b, axis=1) / diag
        return float(np.sqrt(np.mean(e ** 2))), float(e.max())

    forms_out = {}
    for name, r in results.items():
        entry = {
            "solves_per_s": round(r["solves_per_s"], 4),
            "ms_per_step": round(r["ms_per_step"], 3),
            "p50_ms": round(r["p50_ms"], 3),
            "p95_ms": round(r["p95_ms"], 3),
            "p99_ms": round(r["p99_ms"], 3),
            "gn_solves_per_s": round(r["gn_solves_per_s"], 2),
            "pcg_iterations_per_solve": round(r["counters"]["pcg_iterations"] / max(r["counters"]["pcg_launches"], 1),
                                              2),
            "pcg_iterations_most_one_solve": int(r["profiled_counters_max"]),
            "stage_ms_per_profiled_wrap": {k: round(v["ms"], 3) for k, v in r["stages"].items()},
            "roofline_pcg": hbm_roofline(r, "pcg", "pcg_bytes", "pcg_kernel",
                                         PCG_TRAFFIC.get((args.config, name, TOLERANCE))),
            "roofline_assembly": hbm_roofline(r, "assemble", "assemble_bytes", "assembly (rows, node and pair blocks)"),
            "query": query_stats(r),
        }
        if cpu is not None and cpu["vertices"] is not None:
            entry["vertex_rms_vs_oracle"], entry["vertex_max_vs_oracle"] = error(r["vertices"], cpu["vertices"])
        if cpu is not None:
            # Index parity: the finest level's nodes (FPS seeds) and the level
            # sizes, and after a complete wrap the final correspondence
            # counts (accepted and each rejection reason).
            same_shape = r["node_positions"].shape == cpu["node_positions"].shape
            entry["nodes_identical_to_oracle"] = bool(same_shape and np.array_equal(r["node_positions"],
                                                                                    cpu["node_positions"]))
            entry["level_nodes_equal_oracle"] = r["level_nodes"] == cpu["level_nodes"]
            if cpu["correspondence"] is not None:
                counts = ("total", "accepted", "rejected_distance", "rejected_normal", "rejected_slope",
                          "rejected_masked")
                dev = r["report"].get("correspondence", {})
                entry["correspondence_counts_equal_oracle"] = all(dev.get(k) == cpu["correspondence"].get(k)
                                                                  for k in counts)
        forms_out[name] = entry
    if "double" in results:
        for name in results:
            if name == "double":
                continue
            rms, worst = error(results[name]["vertices"], results["double"]["vertices"])
            forms_out[name]["vertex_rms_vs_double"] = rms
            forms_out[name]["vertex_max_vs_double"] = worst

    line = {
        "metric": "wrap solves/s + p50/p99 latency, FP32-refined vs plain FP64, 1×MI355X",
        "value": round(head["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "p50_ms": round(head["p50_ms"], 3),
        "p95_ms": round(head["p95_ms"], 3),
        "p99_ms": round(head["p99_ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if forms[0] == "double" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{job.name}: template {len(job.template_points)} verts / {len(job.template_faces)} tris, "
                        f"target {len(job.target_faces)} tris, {job.nodes_final} nodes, {job.graph_levels} levels x "
                        f"{OUTER} GN steps, 8 graph neighbours, seed {args.seed}",
            "form": forms[0],
            "pcg_tolerance": TOLERANCE,
            "pcg_iteration_cap": MAX_PCG,
            "pcg_iterations_most_one_solve": {n: int(results[n]["profiled_counters_max"]) for n in forms},
        },
        "roofline": hbm_roofline(head, "pcg", "pcg_bytes", "pcg_kernel",
                                 PCG_TRAFFIC.get((args.config, forms[0], TOLERANCE))),
        "cpu_baseline": None if cpu is None else {k: (round(v, 6) if isinstance(v, float) else v)
                                                 for k, v in cpu.items() if k not in CPU_ARRAYS},
        "forms": forms_out,
    }
    if extra_out is not None:
        line.update(extra_out)
    if "double" in results:
        for name in results:
            if name != "double":
                line[f"{name}_over_double"] = round(results[name]["solves_per_s"] / results["double"]["solves_per_s"],
                                                   3)
    if cpu is not None:
        line["gpu_over_cpu_all_cores"] = round(head["solves_per_s"] / cpu["value"], 1)
        line["gpu_over_cpu_1core"] = round(head["solves_per_s"] / cpu["one_core_solves_per_s"], 1)
    if rank == 0:
        text = json.dumps(line)
        print(text, flush=True)
        if args.json_out:
            Path(args.json_out).write_text(text + "\n")
    # The library's cached device memory and events go back before the
    # runtime (and a profiler, when one is attached) shuts down.
    dp.release_cache()
    return 0


if __name__ == "__main__":
    code = main()
    if code:
        sys.exit(code)
