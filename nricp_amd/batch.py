"""Batched uneven instances: a batch of wrap jobs of different sizes run on
one MI355X.

Jobs (as the 64-job batch configuration): humanoid templates of 2k-50k
vertices (log-uniform), targets with 2-3x the template's triangles, nodes ~
vertices / 25, 4 levels x 10 Gauss-Newton steps, FP32 iterations with FP64
refinement, every solve to a 1e-8 relative residual.

Two ways of running the batch are measured:

* packed (nricp_dp.Batch): the jobs in lockstep, one shared launch per stage
  of a damping attempt for every job, one persistent PCG over all jobs'
  block rows, the decisions on the device, each level's ticks replayed from
  graphs built once per level. The autotuner (`tune_packed`) measures the candidate
  shapes on the batch itself and keeps the fastest. A shape is the packed
  PCG's node slots per lane group (its tile: 64 or 128 nodes per workgroup,
  or per level by the single-job session's rule), the jobs per packed batch
  (the launches' width) and the packed batches run at once on host threads
  (one batch's level set-up then overlaps another's ticks);
* concurrent sessions (nricp_dp.Server): each job its own resident session
  on its own streams, `--session-slots` at a time, their kernels sharing the
  device; the process asks HIP for 16 hardware queues (4 by default), so the
  jobs' streams do not serialise on shared queues.

The whole batch is also run one job at a time (the serial baseline), and a
sample of jobs is checked to give the same fitted vertices alone and in the
packed batch (and, for concurrent sessions, alone and in the server).

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


def _packed_chunk(dp, jobs, form: str, pcg_slots: int, graphs: bool) -> tuple[list, dict]:
    batch = dp.Batch(pcg_slots=pcg_slots, use_graphs=graphs)
    totals = {}
    t_add = time.perf_counter()
    for job in jobs:
        a, kw = job_args(job)
        batch.add(*a, precision=form, **kw)
    totals["add_seconds"] = time.perf_counter() - t_add
    results = batch.run()  # releases the GIL while the device works
    totals.update(batch.stats())
    t_close = time.perf_counter()
    batch.close()
    totals["close_seconds"] = time.perf_counter() - t_close
    return results, totals


def run_packed(dp, jobs, form: str, pcg_slots: int, chunk: int, concurrent: int = 1,
               graphs: bool = True) -> tuple[float, list, dict]:
    """The jobs through packed batches of `chunk` jobs each, `concurrent`
    batches at a time on host threads (one batch's level set-up then overlaps
    another's ticks); returns the wall time, the results in job order and the
    summed engine statistics."""
    from concurrent.futures import ThreadPoolExecutor

    chunks = [jobs[s:s + max(1, chunk)] for s in range(0, len(jobs), max(1, chunk))]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=max(1, concurrent)) as pool:
        outs = list(pool.map(lambda c: _packed_chunk(dp, c, form, pcg_slots, graphs), chunks))
    wall = time.perf_counter() - t0
    results, totals = [], {}
    for res, tot in outs:
        results.extend(res)
        for k, v in tot.items():
            totals[k] = totals.get(k, 0.0) + v
    return wall, results, totals


def _key(c) -> str:
    return f"pcg_slots={c[0]},chunk={c[1]},concurrent={c[2]}"


def tune_packed(dp, jobs, form: str, candidates, repeats: int = 1) -> tuple[tuple[int, int, int], dict]:
    """The autotuner: each candidate shape (pcg_slots, jobs per batch,
    batches at once) runs the whole batch `repeats` times; returns the
    fastest and every candidate's median jobs/s."""
    table = {}
    # One untimed run of the whole batch first: the first whole-batch run
    # fills the device-memory cache and the code-object cache for the batch's
    # sizes, which would otherwise be charged to the first candidate.
    run_packed(dp, jobs, form, candidates[0][0], candidates[0][1], candidates[0][2])
    for c in candidates:
        rates = [len(jobs) / run_packed(dp, jobs, form, c[0], c[1], c[2])[0] for _ in range(max(1, repeats))]
        table[_key(c)] = round(float(np.median(rates)), 3)
        log(f"packed: {_key(c)} -> {table[_key(c)]} jobs/s")
    best = max(candidates, key=lambda c: table[_key(c)])
    return best, table


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
    ap.add_argument("--form", default="float")
    ap.add_argument("--candidates", default="2:64:1,2:32:2,2:16:4,2:8:8,1:16:4,0:16:4",
                    help="packed shapes the autotuner measures, pcg_slots:jobs_per_batch:batches_at_once")
    ap.add_argument("--repeats", type=int, default=1, help="runs of the whole batch per candidate")
    ap.add_argument("--session-slots", type=int, default=12,
                    help="jobs on the device at once for the concurrent-session comparison")
    ap.add_argument("--check", type=int, default=4, help="jobs compared alone vs in the batch")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--packed-only", action="store_true", help="measure the packed batch only (profiling runs)")
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
    run_packed(dp, jobs[:8], args.form, 2, 8)

    candidates = [tuple(int(x) for x in c.split(":")) for c in args.candidates.split(",")]
    (pcg_slots, chunk, concurrent), tuning = tune_packed(dp, jobs, args.form, candidates, args.repeats)
    wall, packed, engine = run_packed(dp, jobs, args.form, pcg_slots, chunk, concurrent)
    log(f"packed: {len(jobs) / wall:.1f} jobs/s ({_key((pcg_slots, chunk, concurrent))})")
    alone = [run_packed(dp, [jobs[k]], args.form, pcg_slots, 1)[1][0] for k in range(min(args.check, len(jobs)))]
    same_packed = all(np.array_equal(a["vertices"], packed[k]["vertices"]) for k, a in enumerate(alone))
    log("packed phases (s): " + ", ".join(f"{k} {v:.4f}" for k, v in engine.items()))
    if args.packed_only:
        print(json.dumps({"packed_jobs_per_s": round(len(jobs) / wall, 3), "engine": engine,
                          "tuning_jobs_per_s": tuning, "same_bits_alone_and_in_batch": bool(same_packed)}), flush=True)
        return 0 if same_packed else 3

    session_wall, sessions = run_batch(dp, jobs, args.session_slots, args.form)
    serial_wall, serial = run_batch(dp, jobs, 1, args.form)
    log(f"sessions: {len(jobs) / session_wall:.1f} jobs/s at {args.session_slots} slots, "
        f"{len(jobs) / serial_wall:.1f} jobs/s one at a time")
    same_sessions = all(np.array_equal(sessions[k]["vertices"], serial[k]["vertices"])
                        for k in range(min(args.check, len(jobs))))
    run_ms = np.array([r["run_seconds"] for r in serial]) * 1e3
    diags = [float(np.linalg.norm(j.template_points.max(0) - j.template_points.min(0))) for j in jobs]
    packed_vs_session = max(float(np.sqrt(np.mean(np.sum((p["vertices"] - s["vertices"]) ** 2, axis=1)))) / d
                            for p, s, d in zip(packed, serial, diags))
    line = {
        "metric": "batched uneven instances: jobs/s on one MI355X (whole batch)",
        "workload": f"{len(jobs)} humanoid jobs, {min(sizes)}-{max(sizes)} template vertices, target 2-3x the "
                    f"template's triangles, nodes ~ V/25, 4 levels x {OUTER} GN steps, seed {args.seed}",
        "form": args.form,
        "jobs_per_s": round(len(jobs) / wall, 3),
        "packed": {"pcg_slots": pcg_slots, "jobs_per_batch": chunk, "batches_at_once": concurrent,
                   "jobs_per_s": round(len(jobs) / wall, 3),
                   "batch_seconds": round(wall, 3), "gn_solves_per_s": round(len(jobs) * 4 * OUTER / wall, 1),
                   "engine": {k: round(v, 4) for k, v in engine.items()}, "tuning_jobs_per_s": tuning,
                   "same_bits_alone_and_in_batch": bool(same_packed),
                   "vertex_rms_vs_single_job_sessions_max": packed_vs_session},
        "sessions": {"slots": args.session_slots, "jobs_per_s": round(len(jobs) / session_wall, 3),
                     "same_bits_alone_and_in_server": bool(same_sessions)},
        "serial_jobs_per_s": round(len(jobs) / serial_wall, 3),
        "packed_over_sessions": round(session_wall / wall, 3),
        "packed_over_serial": round(serial_wall / wall, 3),
        "alone_run_ms_p50": round(float(np.percentile(run_ms, 50)), 2),
    }
    text = json.dumps(line)
    print(text, flush=True)
    if args.json_out:
        Path(args.json_out).write_text(text + "\n")
    return 0 if same_packed and same_sessions else 3


if __name__ == "__main__":
    code = main()
    if code:
        sys.exit(code)
