"""Writes tests/golden/large_level_route_a.npz: the oracle's route A (the
FP64 restatement, oracle/src) on a one-level job whose graph has more nodes
than one persistent PCG launch holds on an MI355X (16,384), the case the
device solves with its multi-launch PCG.

Job: geodesic sphere f = 41 (16,812 vertices) onto the bumped, jittered
geodesic sphere f = 60 of workloads.sphere_job (seed 3), one level of 16,500
nodes, 8 graph neighbours, one correspondence update of one Gauss-Newton
step. The oracle takes about a minute here, too long for a GPU test, so the
test reads this fixture. Stored: the fitted vertices' displacement from the
template in FP32 (the test's bound is 1e-5 of the box diagonal), the level
sizes and the final correspondence counts.

Run from the repository root after building the oracle:
    python tests/golden/make_large_level.py
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from nricp_amd import workloads  # noqa: E402

NODES = 16_500
COUNTS = ("total", "accepted", "rejected_distance", "rejected_normal", "rejected_slope", "rejected_masked")
FREQUENCIES = (41, 60)
SEED = 3


def job():
    return workloads.sphere_job(FREQUENCIES[0], FREQUENCIES[1], SEED, nodes_final=NODES, graph_levels=1)


def call_args(j):
    settings = workloads.level_settings(1, outer=1, inner=1)
    kwargs = dict(graph_levels=1, nodes_initial=NODES, growth=2, graph_neighbours=8)
    return (j.template_points, j.template_faces, j.target_points, j.target_faces, settings), kwargs


def main() -> None:
    import nricp_oracle

    j = job()
    args, kwargs = call_args(j)
    r = nricp_oracle.solve_route_a(*args, **kwargs)
    report = json.loads(r["report"])
    np.savez_compressed(
        HERE / "large_level_route_a.npz",
        displacement=(np.asarray(r["vertices"]) - j.template_points).astype(np.float32),
        level_nodes=np.asarray(r["level_nodes"], dtype=np.int64),
        correspondence=np.asarray([report["correspondence"][k] for k in COUNTS], dtype=np.int64),
        outer_iterations=np.asarray([l["outer_iterations"] for l in report["levels"]], dtype=np.int64),
    )


if __name__ == "__main__":
    main()
