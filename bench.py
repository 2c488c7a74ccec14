"""Wrap throughput on one MI355X, beside the FP64 CPU oracle.

A step is one complete wrap of the benchmark job -- device template graph,
target hierarchy, and the whole coarse-to-fine Gauss-Newton schedule -- with
the job's inputs already resident in HBM (the session is created and its
meshes uploaded before the timed region; each step rewinds and re-runs it).

Default workload: the 50k-vertex capsule humanoid onto its posed, bumped,
independently meshed 120k-triangle target with 2% holes; 2,000 graph nodes
over 4 levels, 8 graph neighbours, 10 Gauss-Newton steps per level.

Two device forms are timed: "mixed" (FP32 assembly blocks and PCG iterations,
FP64 right-hand side, energies and refinement; the headline `value`) and
"double" (plain FP64). Both solve each system to a 1e-12 relative residual,
so both stay on the FP64 oracle's trajectory; the error of each against the
oracle's vertices is reported.

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
TOLERANCE = 1e-12
MAX_PCG = 5000
# Device-timed stages of a wrap (HIP events around each stage's launches).
STAGES = ("fps", "graph", "target_bvh", "layout", "deform", "closest_points", "assemble", "pcg", "energy", "local_step")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_job(config: int, seed: int):
    from nricp_amd import workloads

    job = workloads.config_job(config, seed)
    settings = workloads.level_settings(job.graph_levels, outer=10, inner=1)
    kwargs = dict(graph_levels=job.graph_levels, nodes_initial=job.nodes_initial, growth=2, graph_neighbours=8)
    return job, settings, kwargs


def run_form(dp, job, settings, kwargs, form: str, steps: int, warmup: int) -> dict:
    route = "route_a" if form == "double" else "route_b"
    session = dp.RouteBSession(job.template_points, job.template_faces, job.target_points, job.target_faces,
                               settings, route=route, precision="mixed" if form == "mixed" else "double",
                               pcg_tolerance=TOLERANCE, pcg_max_iterations=MAX_PCG, **kwargs)
    for w in range(warmup):
        session.reset()
        session.run()
        log(f"{form}: warmup {w + 1}/{warmup} done")
    session.profile(True)
    before = session.counters()
    stage_before = {s: session.kernel_time(s) for s in STAGES}
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
    stages = {}
    for s, (ms0, n0) in stage_before.items():
        ms1, n1 = session.kernel_time(s)
        stages[s] = {"ms": ms1 - ms0, "launches": n1 - n0}
    delta = {k: after[k] - before.get(k, 0.0) for k in after}
    result = session.result()
    report = json.loads(result["report"])
    return {
        "form": form,
        "solves_per_s": steps / elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "p50_ms": 1e3 * statistics.median(latencies),
        "p99_ms": 1e3 * float(np.percentile(latencies, 99)),
        "gn_solves_per_s": delta["gauss_newton_solves"] / elapsed,
        "counters": delta,
        "stages": stages,
        "vertices": np.asarray(result["vertices"]),
        "report": report,
        "max_query_steps": after.get("query_max_steps", 0.0),
    }


def roofline(form: dict) -> dict:
    pcg = form["stages"]["pcg"]
    seconds = pcg["ms"] * 1e-3
    launches = max(pcg["launches"], 1)
    bytes_per_launch = form["counters"]["pcg_bytes"] / launches
    achieved = (bytes_per_launch / (seconds / launches)) / 1e9 if seconds > 0 else 0.0
    return {
        "kernel": "pcg_kernel",
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": None,
        "avg_launch_us": round(1e6 * seconds / launches, 2),
        "algorithmic_bytes_per_launch": round(bytes_per_launch),
        "iterations_per_launch": round(form["counters"]["pcg_iterations"] / launches, 2),
    }


def query_stats(form: dict) -> dict:
    c = form["counters"]
    rows = max(c.get("query_rows", 0.0), 1.0)
    ms = form["stages"]["closest_points"]["ms"]
    return {
        "queries_per_s": round(rows / (ms * 1e-3), 1) if ms > 0 else None,
        "node_opens_per_query": round(c.get("query_node_opens", 0.0) / rows, 2),
        "triangle_tests_per_query": round(c.get("query_leaf_tests", 0.0) / rows, 2),
        "max_steps_one_query": int(form["max_query_steps"]),
    }


def cpu_baseline(job, settings, kwargs, budget_s: float) -> dict | None:
    import nricp_oracle

    t0 = time.perf_counter()
    result = nricp_oracle.solve_route_a(job.template_points, job.template_faces, job.target_points, job.target_faces,
                                        settings, pcg_tolerance=TOLERANCE, pcg_max_iterations=MAX_PCG, **kwargs)
    seconds = time.perf_counter() - t0
    log(f"cpu oracle: one full wrap in {seconds:.1f} s (budget {budget_s:.0f} s)")
    return {
        "value": 1.0 / seconds,
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": f"one complete wrap of {job.name} (all {job.graph_levels} levels), FP64 route-A restatement, "
                  f"single thread, {seconds:.1f} s",
        "seconds": seconds,
        "vertices": np.asarray(result["vertices"]),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1, help="0: icosphere 10k, 1: humanoid 50k, 2: geodesic 200k")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--forms", default="mixed,double")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--json-out", default=None)
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported by this build "
              "(independent wrap jobs shard across GPUs without collectives; see DESIGN.md)", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    import nricp_amd

    dp = nricp_amd.load()
    info = dp.device_info()
    if not info["available"]:
        print(f"no HIP device: {info['device']}", file=sys.stderr)
        return 1
    log(f"device: {info['device']}")
    t0 = time.perf_counter()
    job, settings, kwargs = build_job(args.config, args.seed)
    log(f"job {job.name}: template {len(job.template_points)} v / {len(job.template_faces)} f, target "
        f"{len(job.target_points)} v / {len(job.target_faces)} f, generated in {time.perf_counter() - t0:.1f} s")

    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    results = {f: run_form(dp, job, settings, kwargs, f, args.steps, args.warmup) for f in forms}
    head = results[forms[0]]

    cpu = None if args.no_cpu or rank != 0 else cpu_baseline(job, settings, kwargs, args.cpu_budget)
    diag = float(np.linalg.norm(job.template_points.max(0) - job.template_points.min(0)))
    forms_out = {}
    for name, r in results.items():
        entry = {
            "solves_per_s": round(r["solves_per_s"], 4),
            "ms_per_step": round(r["ms_per_step"], 3),
            "p50_ms": round(r["p50_ms"], 3),
            "p99_ms": round(r["p99_ms"], 3),
            "gn_solves_per_s": round(r["gn_solves_per_s"], 2),
            "pcg_iterations_per_solve": round(r["counters"]["pcg_iterations"] / max(r["counters"]["pcg_launches"], 1), 2),
            "stage_ms_per_step": {k: round(v["ms"] / args.steps, 3) for k, v in r["stages"].items()},
            "query": query_stats(r),
            "roofline_pcg": roofline(r),
        }
        if cpu is not None:
            err = np.linalg.norm(r["vertices"] - cpu["vertices"], axis=1) / diag
            entry["vertex_rms_vs_oracle"] = float(np.sqrt(np.mean(err ** 2)))
            entry["vertex_max_vs_oracle"] = float(err.max())
        forms_out[name] = entry

    line = {
        "metric": "wrap solves/s + p50/p99 latency, FP32-refined vs plain FP64, 1×MI355X",
        "value": round(head["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "p50_ms": round(head["p50_ms"], 3),
        "p99_ms": round(head["p99_ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if forms[0] == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{job.name}: template {len(job.template_points)} verts / {len(job.template_faces)} tris, "
                        f"target {len(job.target_faces)} tris, {job.nodes_final} nodes, {job.graph_levels} levels x "
                        f"10 GN steps, 8 graph neighbours, seed {args.seed}",
            "form": forms[0],
            "pcg_tolerance": TOLERANCE,
        },
        "roofline": {k: v for k, v in roofline(head).items()},
        "cpu_baseline": None if cpu is None else {k: (round(v, 6) if isinstance(v, float) else v)
                                                 for k, v in cpu.items() if k != "vertices"},
        "forms": forms_out,
    }
    if cpu is not None and "double" in results and "mixed" in results:
        line["mixed_over_double"] = round(results["mixed"]["solves_per_s"] / results["double"]["solves_per_s"], 3)
        line["gpu_over_cpu_1core"] = round(head["solves_per_s"] / cpu["value"], 1)
    if rank == 0:
        text = json.dumps(line)
        print(text, flush=True)
        if args.json_out:
            Path(args.json_out).write_text(text + "\n")
    return 0


if __name__ == "__main__":
    code = main()
    if code:
        sys.exit(code)
