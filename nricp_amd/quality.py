"""Fit quality of a wrap, and acceptance gates against a baseline fit.

The metrics restate the reference's quality harness (bench/geometry.py:247-289):

* correspondence error -- distance from each fitted vertex to where it belongs
  on the target (mean, 95th percentile, max), when the job knows that;
* one-sided chamfer -- mean distance from the fitted vertices to the target
  surface (the device's exact closest-point query);
* stretch -- |edge length / rest length - 1| over the template's unique
  edges (mean and 95th percentile);
* foldover -- the fraction of template faces whose normal turned against the
  rest template's (the template is given pre-aligned, so the rest template is
  the aligned one).

Gates (bench/geometry.py:292-341, tolerances from bench/config/bench_config.json):
a result fails when its corr_mean exceeds 1.05 x the baseline's, its
corr_p95 1.1 x, its stretch_p95 the baseline's + 0.05, or its foldover the
baseline's + 0.01; a non-finite metric always fails. The reference anchors the
gates on its best recorded method per pair; here the baseline is the plain
FP64 route (route A semantics, FP64 throughout) on the same job, and the
tests also gate the device forms against the FP64 CPU route.

Run from the repository root:  python -m nricp_amd.quality --config 1
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
for sub in ("", "nricp_amd"):
    p = str(ROOT / sub) if sub else str(ROOT)
    if p not in sys.path:
        sys.path.insert(0, p)

GATES = {"corr_mean_ratio": 1.05, "corr_p95_ratio": 1.1, "stretch_p95_margin": 0.05, "foldover_margin": 0.01}


def unique_edges(faces: np.ndarray) -> np.ndarray:
    f = np.asarray(faces, dtype=np.int64)
    edges = np.sort(np.vstack((f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]])), axis=1)
    return np.unique(edges, axis=0)


def face_normals(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:
    t = np.asarray(vertices)[np.asarray(faces)]
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    return n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-300)


def correspondence_errors(result: np.ndarray, truth: np.ndarray) -> dict:
    d = np.linalg.norm(np.asarray(result) - np.asarray(truth), axis=1)
    return {"corr_mean": float(d.mean()), "corr_p95": float(np.percentile(d, 95)), "corr_max": float(d.max())}


def stretch(result: np.ndarray, rest: np.ndarray, faces: np.ndarray) -> tuple[float, float]:
    e = unique_edges(faces)
    r = np.linalg.norm(rest[e[:, 0]] - rest[e[:, 1]], axis=1)
    c = np.linalg.norm(result[e[:, 0]] - result[e[:, 1]], axis=1)
    rel = np.abs(c / np.maximum(r, 1e-12) - 1.0)
    return float(rel.mean()), float(np.percentile(rel, 95))


def foldover(result: np.ndarray, rest: np.ndarray, faces: np.ndarray) -> float:
    a, b = face_normals(rest, faces), face_normals(result, faces)
    return float((np.einsum("ij,ij->i", a, b) < 0).mean())


def one_sided_chamfer(result: np.ndarray, target_points: np.ndarray, target_faces: np.ndarray, closest) -> float:
    """Mean distance from the fitted vertices to the target surface;
    ``closest`` is a closest_points(target_points, target_faces, queries)."""
    out = closest(np.ascontiguousarray(target_points, dtype=np.float64),
                  np.ascontiguousarray(target_faces, dtype=np.int64),
                  np.ascontiguousarray(result, dtype=np.float64))
    return float(np.asarray(out["distances"]).mean())


def evaluate(result: np.ndarray, job, closest) -> dict:
    result = np.asarray(result, dtype=np.float64)
    diag = float(np.linalg.norm(job.template_points.max(0) - job.template_points.min(0)))
    m = {}
    if job.ground_truth is not None:
        m.update(correspondence_errors(result, job.ground_truth))
        m["corr_mean_fraction_of_diagonal"] = m["corr_mean"] / diag
    m["chamfer_A2B"] = one_sided_chamfer(result, job.target_points, job.target_faces, closest)
    m["stretch_mean"], m["stretch_p95"] = stretch(result, job.template_points, job.template_faces)
    m["foldover"] = foldover(result, job.template_points, job.template_faces)
    return m


def gate_failures(metrics: dict, baseline: dict) -> list[str]:
    """Which gates `metrics` violates against `baseline` (empty: passes)."""
    limits = [("stretch_p95", baseline["stretch_p95"] + GATES["stretch_p95_margin"]),
              ("foldover", baseline["foldover"] + GATES["foldover_margin"])]
    if "corr_mean" in baseline and "corr_mean" in metrics:
        limits = [("corr_mean", baseline["corr_mean"] * GATES["corr_mean_ratio"]),
                  ("corr_p95", baseline["corr_p95"] * GATES["corr_p95_ratio"])] + limits
    failures = []
    for name, limit in limits:
        value = metrics[name]
        if not math.isfinite(value):
            failures.append(f"{name} is not finite ({value!r})")
        elif value > limit:
            failures.append(f"{name} {value:.6g} > {limit:.6g} (baseline {baseline[name]:.6g})")
    return failures


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--config", type=int, default=1, help="0: icosphere 10k, 1: humanoid 50k, 2: geodesic 200k, 3: humanoid 200k")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--forms", default="mixed,double")
    ap.add_argument("--json-out", default=None)
    args = ap.parse_args(argv)

    from nricp_amd import workloads
    import nricp_amd

    job = workloads.config_job(args.config, args.seed)
    settings = workloads.level_settings(job.graph_levels, outer=10, inner=1)
    kwargs = dict(graph_levels=job.graph_levels, nodes_initial=job.nodes_initial, growth=2, graph_neighbours=8,
                  pcg_tolerance=1e-12, pcg_max_iterations=5000)
    args_ = (job.template_points, job.template_faces, job.target_points, job.target_faces, settings)
    dp = nricp_amd.load()
    t0 = time.perf_counter()
    base = dp.solve_route_a(*args_, **kwargs)
    base_seconds = time.perf_counter() - t0
    out = {"workload": job.name, "baseline": "plain FP64 route (route_a)", "baseline_seconds": round(base_seconds, 3)}
    out["baseline_metrics"] = evaluate(base["vertices"], job, dp.closest_points)
    out["forms"] = {}
    for form in [f.strip() for f in args.forms.split(",") if f.strip()]:
        t0 = time.perf_counter()
        res = dp.solve_route_b(*args_, precision=form, **kwargs)
        seconds = time.perf_counter() - t0
        m = evaluate(res["vertices"], job, dp.closest_points)
        out["forms"][form] = {"metrics": m, "gate_failures": gate_failures(m, out["baseline_metrics"]),
                              "seconds": round(seconds, 3)}
    text = json.dumps(out)
    print(text, flush=True)
    if args.json_out:
        Path(args.json_out).write_text(text + "\n")
    return 0 if all(not f["gate_failures"] for f in out["forms"].values()) else 3


if __name__ == "__main__":
    code = main()
    if code:
        sys.exit(code)
