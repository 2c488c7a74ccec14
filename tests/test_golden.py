"""The oracle (CPU) and the device path (GPU) against fixtures written by the
reference's own CPU code (tests/golden/make_golden.py over the reference's
Eigen-free translation units).

Index results must be identical. FP64 results computed in the reference's
operation order must be identical to the last bit (normals, seeds, bindings,
closest points on a mesh with no equidistant ties, filters, robust weights,
the report text); reductions whose order differs (the assembled system) are
held to 1e-12 relative.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def fixture(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_vertex_normals(impl):
    g = fixture("vertex_normals")
    out = np.asarray(impl.vertex_normals(g["points"], g["faces"]))
    assert np.array_equal(out, g["normals"])


def test_clean_mesh(host_impl):
    g = fixture("clean_mesh")
    r = host_impl.clean_mesh(g["points"], g["faces"])
    assert np.array_equal(np.asarray(r["points"]), g["clean_points"])
    assert np.array_equal(np.asarray(r["faces"]), g["clean_faces"])
    assert [r["duplicate_faces"], r["degenerate_faces"], r["isolated_vertices"]] == list(g["counts"])


def test_normalization(host_impl):
    g = fixture("normalization")
    r = host_impl.normalization(g["points"])
    assert np.array_equal(np.asarray(r["center"]).reshape(3), g["center"])
    assert float(r["scale"]) == float(g["scale"][0])


@pytest.mark.parametrize("first", [0, 7])
def test_farthest_point_seeds(impl, first):
    g = fixture("farthest_point_seeds")
    out = np.asarray(impl.farthest_point_seeds(g["points"], int(g["count"]), first))
    assert np.array_equal(out, g[f"first_{first}"])


@pytest.mark.parametrize("tag", ["many", "few", "three"])
def test_bind_vertices(impl, tag):
    g = fixture("bind_vertices")
    r = impl.bind_vertices(g["points"], g[f"{tag}_nodes"], int(g[f"{tag}_k"]))
    counts = np.asarray(r["counts"])
    assert np.array_equal(counts, g[f"{tag}_counts"])
    nodes, weights = np.asarray(r["nodes"]), np.asarray(r["weights"])
    for v, c in enumerate(counts):
        assert np.array_equal(nodes[v, :c], g[f"{tag}_bound"][v, :c])
        assert np.array_equal(weights[v, :c], g[f"{tag}_weights"][v, :c])


@pytest.mark.parametrize("k", [4, 8])
def test_build_node_edges(impl, k):
    g = fixture("build_node_edges")
    out = np.asarray(impl.build_node_edges(g["nodes"], k))
    assert np.array_equal(out, g[f"edges_{k}"])


def test_closest_points(impl):
    """Identical answers, with one documented freedom: when two triangles are
    at exactly the same squared distance, the reference keeps the one its
    hierarchy visits first and the device the lower triangle index. The
    distance is then the same bits; the closest points are the same feature
    (a shared vertex, or a shared edge evaluated from its two ends) and agree
    to a few ulps."""
    g = fixture("closest_points")
    r = impl.closest_points(g["target_points"], g["target_faces"], g["queries"])
    tri = np.asarray(r["triangles"])
    pts, nrm = np.asarray(r["points"]), np.asarray(r["normals"])
    assert np.array_equal(np.asarray(r["distances"]), g["distances"])
    same = tri == g["triangles"]
    if impl.__name__ == "nricp_oracle":
        assert same.all()
    assert np.array_equal(pts[same], g["points"][same])
    assert np.array_equal(nrm[same], g["normals"][same])
    faces = g["target_faces"]
    for i in np.nonzero(~same)[0]:
        assert set(faces[tri[i]]) & set(faces[g["triangles"][i]]), i  # neighbours sharing the tied feature
    np.testing.assert_allclose(pts[~same], g["points"][~same], rtol=0, atol=1e-15)
    np.testing.assert_allclose(nrm[~same], g["normals"][~same], rtol=0, atol=1e-12)


@pytest.mark.parametrize("tag", ["plain", "filtered"])
def test_find_correspondences(impl, tag):
    g = fixture("find_correspondences")
    dist, mincos, slope, use_mask = g[f"{tag}_settings"]
    r = impl.find_correspondences(g["target_points"], g["target_faces"], g["queries"], g["query_normals"],
                                  distance_threshold=float(dist), min_normal_cosine=float(mincos),
                                  min_slope=float(slope), mask=g["mask"] if use_mask else None)
    assert np.array_equal(np.asarray(r["rejection"]), g[f"{tag}_rejection"])
    c = r["counts"]
    got = [c["total"], c["accepted"], c["rejected_distance"], c["rejected_normal"], c["rejected_slope"],
           c["rejected_masked"]]
    assert got == list(g[f"{tag}_counts"])
    stats = np.array([c["distance_mean"], c["distance_p95"], c["distance_max"]])
    np.testing.assert_allclose(stats, g[f"{tag}_stats"], rtol=1e-14, atol=0)


@pytest.mark.parametrize("kernel", ["huber", "geman_mcclure"])
@pytest.mark.parametrize("threshold", [0.0, 0.05, 0.5])
def test_robust_weights(impl, kernel, threshold):
    g = fixture("robust_weights")
    out = np.asarray(impl.robust_weights(kernel, g["squared"], threshold))
    assert np.array_equal(out, g[f"{kernel}_{threshold}"])


def test_assemble_normal_equations(host_impl):
    g = fixture("assemble_normal_equations")
    r = host_impl.assemble_normal_equations(g["kind"], g["rows"], g["values"], g["slots"], g["nodes"],
                                            g["jacobians"], int(g["node_count"]))
    scale = np.abs(g["diagonal"]).max()
    np.testing.assert_allclose(np.asarray(r["diagonal"]).reshape(g["diagonal"].shape), g["diagonal"], rtol=0,
                               atol=1e-12 * scale)
    np.testing.assert_allclose(np.asarray(r["rhs"]).reshape(g["rhs"].shape), g["rhs"], rtol=0,
                               atol=1e-12 * np.abs(g["rhs"]).max())
    assert np.array_equal(np.asarray(r["edges"]), g["pairs"])
    np.testing.assert_allclose(np.asarray(r["edge_blocks"]).reshape(g["blocks"].shape), g["blocks"], rtol=0,
                               atol=1e-12 * scale)


@pytest.mark.parametrize("lam,scaling", [(0.0, 0), (0.3, 1), (0.3, 0)])
def test_damped_solve_matches_the_reference_damping(impl, lam, scaling):
    """The reference's damped diagonal, solved densely, against the solve of
    the undamped system with the same damping applied by the implementation."""
    g = fixture("assemble_normal_equations")
    n = int(g["node_count"])
    damped = g[f"damped_{lam}_{scaling}"]
    dense = np.zeros((6 * n, 6 * n))
    for j in range(n):
        dense[6 * j:6 * j + 6, 6 * j:6 * j + 6] = damped[j]
    for (a, b), blk in zip(g["pairs"], g["blocks"]):
        dense[6 * a:6 * a + 6, 6 * b:6 * b + 6] = blk
        dense[6 * b:6 * b + 6, 6 * a:6 * a + 6] = blk.T
    expected = np.linalg.solve(dense, g["rhs"].reshape(-1))
    kwargs = {"tolerance": 1e-13} if impl.__name__ == "nricp_dp" else {}
    r = impl.solve_normal_equations(g["diagonal"], g["pairs"], g["blocks"], g["rhs"], damping=lam,
                                    diagonal_scaling=bool(scaling), **kwargs)
    assert r["ok"]
    x = np.asarray(r["delta"]).reshape(-1)
    np.testing.assert_allclose(x, expected, rtol=0, atol=1e-9 * np.abs(expected).max())


def _report_args(case):
    energies = [float(e) for e in case["energies"]]
    names = ["point_to_point", "point_to_plane", "smoothness", "anchor", "landmark"]
    cnames = ["total", "accepted", "rejected_distance", "rejected_normal", "rejected_slope", "rejected_masked"]
    corr = dict(zip(cnames, case["counts"]))
    corr.update(distance_mean=case["stats"][0], distance_p95=case["stats"][1], distance_max=case["stats"][2])
    levels = [{"level": l, "nodes": 100 * (l + 1), "outer_iterations": 3 + l, "inner_iterations": 7 + l,
               "energy_start": 1.0 / (l + 3.0), "energy_end": 1.0 / (l + 7.0), "correspondence_mean": 0.125 * l,
               "converged": l % 2 == 0} for l in range(case["levels"])]
    settings = {f"level{l}_smoothness": 0.1 * (l + 1) for l in range(case["levels"])}
    return ("route_a", "double", 0.25, 1.75, dict(zip(names, energies)), corr, levels, settings)


@pytest.mark.parametrize("index", [0, 1])
def test_report_text(host_impl, index):
    case = json.loads((GOLDEN / "report.json").read_text())[index]
    assert host_impl.serialize_report(*_report_args(case)) == case["json"]
