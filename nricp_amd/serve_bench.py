"""Serving benchmark: wrap jobs arriving as a Poisson process, served by
nricp_dp.Server on one MI355X.

Jobs: a pool of uneven jobs drawn like the 64-job batch configuration
(humanoid templates of 2k-50k vertices, log-uniform; targets with 2-3x the
template's triangles; nodes ~ vertices / 25; 4 levels x 10 Gauss-Newton
steps; every solve to a 1e-8 relative residual), plus the 200k-vertex job, which 5 % of the arrivals request. At each
offered load (jobs/s) a fixed number of arrivals is drawn with exponential
gaps; a submitter thread submits each job at its arrival time. Latency is a
job's submission-to-result time as the server measured it (queueing
included). Throughput is the jobs completed over the span from the first
arrival to the last completion.
The process asks HIP for 16 hardware queues (4 by default), so concurrent
jobs' streams do not serialise on shared queues.

Run from the repository root:  python -m nricp_amd.serve_bench
Prints one JSON line per (form, offered load), then a summary line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import threading
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
    print(f"[serve] {msg}", file=sys.stderr, flush=True)


def _make_job(k: int, seed: int):
    return workloads.batch_jobs(1, seed=seed + 7919 * k, kind="humanoid")[0]


def build_pool(size: int, seed: int, procs: int) -> list:
    """The pool of uneven jobs, generated in worker processes (before the
    GPU is touched, so the workers start from a process with no device
    state)."""
    if procs > 1:
        import multiprocessing as mp

        with mp.get_context("spawn").Pool(procs) as pool:
            return pool.starmap(_make_job, [(k, seed) for k in range(size)])
    return [_make_job(k, seed) for k in range(size)]


def job_args(job):
    settings = workloads.level_settings(job.graph_levels, outer=OUTER, inner=1)
    return ((job.template_points, job.template_faces, job.target_points, job.target_faces, settings),
            dict(KWARGS, graph_levels=job.graph_levels, nodes_initial=job.nodes_initial))


def run_load(server, pool, big, rate: float, arrivals: int, form: str, rng: np.random.Generator) -> dict:
    gaps = rng.exponential(1.0 / rate, size=arrivals)
    times = np.cumsum(gaps) - gaps[0]
    picks = [big if rng.random() < 0.05 else pool[int(rng.integers(len(pool)))] for _ in range(arrivals)]
    tickets: list = [None] * arrivals
    submitted = threading.Event()
    start = time.perf_counter()

    def submitter():
        for i, (t, job) in enumerate(zip(times, picks)):
            delay = start + t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            a, kw = job_args(job)
            tickets[i] = server.submit(*a, precision=form, **kw)
        submitted.set()

    th = threading.Thread(target=submitter)
    th.start()
    results = []
    for i in range(arrivals):
        while tickets[i] is None:
            time.sleep(0.0005)
        results.append(server.result(tickets[i]))
    th.join()
    end = time.perf_counter()
    latency = np.array([r["queued_seconds"] + r["run_seconds"] for r in results]) * 1e3
    queued = np.array([r["queued_seconds"] for r in results]) * 1e3
    sizes = [len(j.template_points) for j in picks]
    return {
        "form": form,
        "offered_jobs_per_s": rate,
        "arrivals": arrivals,
        "solves_per_s": round(arrivals / (end - start), 3),
        "p50_ms": round(float(np.percentile(latency, 50)), 2),
        "p95_ms": round(float(np.percentile(latency, 95)), 2),
        "p99_ms": round(float(np.percentile(latency, 99)), 2),
        "mean_queue_ms": round(float(queued.mean()), 2),
        "mean_template_vertices": int(np.mean(sizes)),
        "large_jobs": int(sum(1 for j in picks if j is big)),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--rates", default="5,10,20,50,100,200")
    ap.add_argument("--arrivals", type=int, default=40, help="jobs per offered load")
    ap.add_argument("--pool", type=int, default=16, help="distinct uneven jobs to draw from")
    ap.add_argument("--max-active", type=int, default=4)
    ap.add_argument("--forms", default="mixed,double")
    ap.add_argument("--seed", type=int, default=3000)
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="hardware queues HIP gives the process (GPU_MAX_HW_QUEUES; unless already set): with the "
                         "default 4, the jobs' streams share queues and their kernels serialise")
    args = ap.parse_args(argv)
    # Read once, when HIP initialises: set before anything touches the device.
    os.environ.setdefault("GPU_MAX_HW_QUEUES", str(args.hw_queues))

    t0 = time.perf_counter()
    procs = min(len(os.sched_getaffinity(0)), 16)
    pool = build_pool(args.pool, args.seed, procs)
    big = workloads.config_job(2, 0)
    log(f"pool of {len(pool)} jobs ({min(len(j.template_points) for j in pool)}-"
        f"{max(len(j.template_points) for j in pool)} template vertices) + the 200k job, generated in "
        f"{time.perf_counter() - t0:.1f} s")

    import nricp_amd

    dp = nricp_amd.load()
    info = dp.device_info()
    if not info["available"]:
        print(f"no HIP device: {info['device']}", file=sys.stderr)
        return 1
    lines = []
    for form in [f.strip() for f in args.forms.split(",") if f.strip()]:
        with dp.Server(max_active=args.max_active) as server:
            # Warm-up: every pool job once, and the large one.
            for job in pool + [big]:
                a, kw = job_args(job)
                server.result(server.submit(*a, precision=form, **kw))
            for k, rate in enumerate(float(r) for r in args.rates.split(",")):
                # The same arrivals (gaps and jobs) for every form at a given load.
                rng = np.random.default_rng([args.seed, k])
                line = run_load(server, pool, big, rate, args.arrivals, form, rng)
                line["max_active"] = args.max_active
                print(json.dumps(line), flush=True)
                lines.append(line)
    summary = {
        "metric": "serving: p50/p95/p99 latency and solves/s vs offered load, 1xMI355X",
        "workload": f"Poisson arrivals; {args.pool} uneven humanoid jobs (2k-50k template vertices, target 2-3x the "
                    f"template's triangles, nodes ~ V/25, 4 levels x {OUTER} GN steps) + 5 % 200k-vertex jobs",
        "max_active": args.max_active,
        "loads": lines,
    }
    print(json.dumps(summary), flush=True)
    if args.json_out:
        Path(args.json_out).write_text(json.dumps(summary) + "\n")
    return 0


if __name__ == "__main__":
    code = main()
    if code:
        sys.exit(code)
