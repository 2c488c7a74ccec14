"""The assembled normal equations and their solve, as in the reference's
tests/test_normal_equations.py: the assembly against an independent numpy
assembly from the block Jacobians, the dense expansion, the solve against a
dense solve, the descent direction (the rhs sign), damping, and the fused
single-pass system against the block path."""

from __future__ import annotations

import numpy as np
import pytest

from meshes import all_weights, energy_args, graph_state, sphere_patch


def system(mod, pts, state, weights=None, mask=None):
    r = mod.energy_blocks(*energy_args(pts, state), weights or all_weights(), mask)
    b = r["blocks"]
    s = mod.assemble_normal_equations(b["kind"], b["rows"], b["values"], b["slots"], b["nodes"], b["jacobians"],
                                      len(state["nodes"]))
    return r, b, s


def numpy_system(b, n):
    rows, values, slots = np.asarray(b["rows"]), np.asarray(b["values"]), np.asarray(b["slots"])
    nodes, jac = np.asarray(b["nodes"]), np.asarray(b["jacobians"])
    diag, rhs, pairs = np.zeros((n, 6, 6)), np.zeros((n, 6)), {}
    for i in range(len(rows)):
        for s in range(slots[i]):
            a = nodes[i, s]
            ja, r = jac[i, s, :rows[i]], values[i, :rows[i]]
            diag[a] += ja.T @ ja
            rhs[a] -= ja.T @ r
            for t in range(s + 1, slots[i]):
                c = nodes[i, t]
                jc = jac[i, t, :rows[i]]
                key, blk = ((a, c), ja.T @ jc) if a < c else ((c, a), jc.T @ ja)
                pairs[key] = pairs.get(key, 0.0) + blk
    return diag, rhs, pairs


def test_assembly_equals_an_independent_assembly(host_impl, oracle):
    pts, _ = sphere_patch()
    state = graph_state(oracle, pts, 13, 10)
    _, b, s = system(host_impl, pts, state)
    diag, rhs, pairs = numpy_system(b, 13)
    np.testing.assert_allclose(np.asarray(s["diagonal"]), diag, rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.asarray(s["rhs"]), rhs, rtol=0, atol=1e-12)
    edges = [tuple(e) for e in np.asarray(s["edges"])]
    assert sorted(edges) == sorted(pairs) and edges == sorted(edges)
    for e, blk in zip(edges, np.asarray(s["edge_blocks"])):
        np.testing.assert_allclose(blk, pairs[e], rtol=0, atol=1e-12)


def test_dense_expansion_is_the_symmetric_block_matrix(host_impl, oracle):
    pts, _ = sphere_patch()
    _, b, s = system(host_impl, pts, graph_state(oracle, pts, 13, 11))
    dense = np.asarray(host_impl.normal_equations_dense(s["diagonal"], s["edges"], s["edge_blocks"], s["rhs"]))
    diag, _, pairs = numpy_system(b, 13)
    expect = np.zeros((78, 78))
    for j in range(13):
        expect[6 * j:6 * j + 6, 6 * j:6 * j + 6] = diag[j]
    for (a, c), blk in pairs.items():
        expect[6 * a:6 * a + 6, 6 * c:6 * c + 6] += blk
        expect[6 * c:6 * c + 6, 6 * a:6 * a + 6] += blk.T
    np.testing.assert_allclose(dense, expect, rtol=0, atol=1e-12)
    np.testing.assert_allclose(dense, dense.T, rtol=0, atol=1e-12)


def test_the_solve_matches_a_dense_solve(impl, oracle):
    pts, _ = sphere_patch()
    _, _, s = system(oracle, pts, graph_state(oracle, pts, 13, 12))
    dense = np.asarray(oracle.normal_equations_dense(s["diagonal"], s["edges"], s["edge_blocks"], s["rhs"]))
    rhs = np.asarray(s["rhs"]).reshape(-1)
    expected = np.linalg.solve(dense + 1e-3 * np.eye(78), rhs)
    r = impl.solve_normal_equations(s["diagonal"], s["edges"], s["edge_blocks"], s["rhs"], damping=1e-3)
    assert r["ok"]
    np.testing.assert_allclose(np.asarray(r["delta"]).reshape(-1), expected, rtol=0, atol=1e-9)
    assert r["residual_norm"] < 1e-8 * max(np.linalg.norm(rhs), 1.0)


def test_the_step_is_a_descent_direction(impl, oracle):
    pts, _ = sphere_patch()
    state = graph_state(oracle, pts, 13, 13)
    r0, _, s = system(oracle, pts, state)
    delta = np.asarray(impl.solve_normal_equations(s["diagonal"], s["edges"], s["edge_blocks"], s["rhs"],
                                                   damping=1e-6)["delta"])
    moved = impl.retract_deformation(state["rotations"], state["translations"], state["nodes"], delta)
    after = oracle.energy_blocks(*energy_args(pts, dict(state, rotations=moved["rotations"],
                                                        translations=moved["translations"])), all_weights(), None)
    assert after["energies"]["total"] < r0["energies"]["total"]


def test_more_damping_means_a_shorter_step(impl, oracle):
    pts, _ = sphere_patch()
    _, _, s = system(oracle, pts, graph_state(oracle, pts, 13, 14))
    lengths = [np.linalg.norm(np.asarray(impl.solve_normal_equations(s["diagonal"], s["edges"], s["edge_blocks"],
                                                                     s["rhs"], damping=lam)["delta"]))
               for lam in (1e-6, 1e-2, 1.0)]
    assert lengths[0] > lengths[1] > lengths[2]


def test_scaled_damping_solves_a_system_without_the_anchor(impl, oracle):
    pts, _ = sphere_patch()
    w = dict(all_weights(), anchor=0.0)
    _, _, s = system(oracle, pts, graph_state(oracle, pts, 13, 15), weights=w)
    plain = impl.solve_normal_equations(s["diagonal"], s["edges"], s["edge_blocks"], s["rhs"], damping=1e-4)
    scaled = impl.solve_normal_equations(s["diagonal"], s["edges"], s["edge_blocks"], s["rhs"], damping=1e-4,
                                         diagonal_scaling=True)
    assert np.isfinite(np.asarray(plain["delta"])).all() and np.isfinite(np.asarray(scaled["delta"])).all()
    assert scaled["ok"]


@pytest.mark.parametrize("masked", [False, True])
def test_the_fused_system_reproduces_the_block_path(impl, oracle, masked):
    pts, _ = sphere_patch()
    state = graph_state(oracle, pts, 13, 21)
    mask = None
    if masked:
        mask = np.ones(len(pts), dtype=np.int64)
        mask[[2, 5, 6]] = 0
    ref, _, s = system(oracle, pts, state, mask=mask)
    f = impl.fused_system(*energy_args(pts, state), all_weights(), mask)
    scale = np.abs(np.asarray(s["diagonal"])).max()
    np.testing.assert_allclose(np.asarray(f["diagonal"]), np.asarray(s["diagonal"]), rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(np.asarray(f["rhs"]), np.asarray(s["rhs"]), rtol=0, atol=1e-12 * scale)
    assert np.array_equal(np.asarray(f["edges"]), np.asarray(s["edges"]))
    np.testing.assert_allclose(np.asarray(f["edge_blocks"]), np.asarray(s["edge_blocks"]), rtol=0,
                               atol=1e-12 * scale)
    for k, v in ref["energies"].items():
        assert f["energies"][k] == pytest.approx(v, rel=1e-12)
