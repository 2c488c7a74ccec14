"""Batched uneven instances: a batch of wrap jobs of different sizes run on
one MI355X, with the number of jobs on the device at once chosen by a
measured sweep.

Jobs (as the 64-job batch configuration): humanoid templates of 2k-50k
vertices (log-uniform), targets with 2-3x the template's triangles, nodes ~
vertices / 25, 4 levels x 10 Gauss-Newton steps, FP32 iterations with FP64
refinement, every solve to a 1e-8 relative residual. The batch runs through nricp_dp.Server, so each job is its own
resident session on its own streams and the jobs' kernels share the device:
while one job waits on a grid barrier or a host decision, the others' work
fills the compute units. The process asks HIP for 16 hardware queues (4 by
default), so the jobs' streams do not serialise on shared queues.

Tuning: the whole batch is run `--repeats` times at each candidate slot
count (a run takes about a second; a subset of the jobs tunes poorly, since
the mix of sizes decides how many jobs overlap well). The count with the best
median jobs/s then runs the batch once more for the reported figure. The
whole batch is also run one job at a time (the serial baseline), and a sample
of jobs is checked to give the same fitted vertices alone and in the batch.

Run from the repository root:  python -m nricp_amd.batch
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
for sub in ("", "nricp_amd"):
    p = str(ROOT / sub) if sub else str(ROOT)
    if p not in sys.path:
        sys.path.insert(0, p)

from nricp_amd import workloads  # noqa: E402

OUTER = 10
KWARGS = dict(growth=2, graph_neighbours=8, pcg_tolerance=1e-8, pcg_max_iterations=5000)


def log(msg: str) -> None:
    print(f"[batch] {msg}", file=sys.stderr, flush=True)


def _make_job(k: int, count: int, seed: int):
    return workloads.batch_jobs(count, seed=seed, kind="humanoid", indices=[k])[0]


def make_batch(count: int, seed: int) -> list:
    """The batch (workloads.batch_jobs: sizes log-uniform, job k seeded
    seed + k), generated in worker processes before the GPU is touched."""
    procs = min(len(os.sched_getaffinity(0)), 16)
    if procs > 1:
        import multiprocessing as mp

        with mp.get_context("spawn").Pool(procs) as pool:
            return pool.starmap(_make_job, [(k, count, seed) for k in range(count)])
    return [_make_job(k, count, seed) for k in range(count)]


def job_args(job):
    settings = workloads.level_settings(job.graph_levels, outer=OUTER, inner=1)
    return ((job.template_points, job.template_faces, job.target_points, job.target_faces, settings),
            dict(KWARGS, graph_levels=job.graph_levels, nodes_initial=job.nodes_initial))


def run_batch(dp, jobs, slots: int, form: str) -> tuple[float, list]:
    """All jobs submitted at once to a server with `slots` slots; returns the
    wall time and the results in job order."""
    with dp.Server(max_active=slots) as server:
        t0 = time.perf_counter()
        tickets = []
        for job in jobs:
            a, kw = job_args(job)
            tickets.append(server.submit(*a, precision=form, **kw))
        results = [server.result(t) for t in tickets]
        return time.perf_counter() - t0, results


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--form", default="mixed")
    ap.add_argument("--candidates", default="1,2,4,6,8,12,16", help="slot counts the tuner measures")
    ap.add_argument("--repeats", type=int, default=2, help="runs of the whole batch per candidate")
    ap.add_argument("--check", type=int, default=4, help="jobs compared alone vs in the batch")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="hardware queues HIP gives the process (GPU_MAX_HW_QUEUES; unless already set): with the "
                         "default 4, the jobs' streams share queues and their kernels serialise")
    args = ap.parse_args(argv)
    # Read once, when HIP initialises: set before anything touches the device.
    os.environ.setdefault("GPU_MAX_HW_QUEUES", str(args.hw_queues))

    t0 = time.perf_counter()
    jobs = make_batch(args.jobs, args.seed)
    sizes = [len(j.template_points) for j in jobs]
    log(f"{len(jobs)} jobs, {min(sizes)}-{max(sizes)} template vertices ({sum(sizes)} in all), generated in "
        f"{time.perf_counter() - t0:.1f} s")

    import nricp_amd

    dp = nricp_amd.load()
    info = dp.device_info()
    if not info["available"]:
        print(f"no HIP device: {info['device']}", file=sys.stderr)
        return 1
    # Warm-up: code objects loaded, allocator cache filled.
    run_batch(dp, jobs[:8], 4, args.form)

    tuning = {}
    for slots in [int(c) for c in args.candidates.split(",")]:
        rates = [len(jobs) / run_batch(dp, jobs, slots, args.form)[0] for _ in range(max(1, args.repeats))]
        tuning[slots] = round(float(np.median(rates)), 3)
        log(f"tuning: {slots} slots -> {tuning[slots]} jobs/s (median of {len(rates)})")
    best = max(tuning, key=lambda k: tuning[k])

    wall, results = run_batch(dp, jobs, best, args.form)
    serial_wall, serial = run_batch(dp, jobs, 1, args.form)
    latency = np.array([r["queued_seconds"] + r["run_seconds"] for r in results]) * 1e3
    run_ms = np.array([r["run_seconds"] for r in serial]) * 1e3
    same = all(np.array_equal(results[k]["vertices"], serial[k]["vertices"]) for k in range(min(args.check, len(jobs))))
    line = {
        "metric": "batched uneven instances: jobs/s on one MI355X (whole batch)",
        "workload": f"{len(jobs)} humanoid jobs, {min(sizes)}-{max(sizes)} template vertices, target 2-3x the "
                    f"template's triangles, nodes ~ V/25, 4 levels x {OUTER} GN steps, seed {args.seed}",
        "form": args.form,
        "slots": best,
        "jobs_per_s": round(len(jobs) / wall, 3),
        "batch_seconds": round(wall, 3),
        "serial_jobs_per_s": round(len(jobs) / serial_wall, 3),
        "speedup_over_serial": round(serial_wall / wall, 3),
        "gn_solves_per_s": round(len(jobs) * 4 * OUTER / wall, 1),
        "latency_p50_ms": round(float(np.percentile(latency, 50)), 2),
        "latency_p99_ms": round(float(np.percentile(latency, 99)), 2),
        "alone_run_ms_p50": round(float(np.percentile(run_ms, 50)), 2),
        "tuning_jobs_per_s": tuning,
        "same_bits_alone_and_in_batch": bool(same),
    }
    text = json.dumps(line)
    print(text, flush=True)
    if args.json_out:
        Path(args.json_out).write_text(text + "\n")
    return 0 if same else 3


if __name__ == "__main__":
    code = main()
    if code:
        sys.exit(code)
