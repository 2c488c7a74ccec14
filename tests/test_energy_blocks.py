"""Residual blocks: every kind present, energies equal to their rows, the mask
removes exactly the data rows, analytic Jacobians against central differences
of the actual residual function, and the robust kernels' closed forms -- the
checks of the reference's tests/test_energy_blocks.py, restated."""

from __future__ import annotations

import numpy as np
import pytest

from meshes import all_weights, energy_args, graph_state, sphere_patch

P2P, PLANE, SMOOTH, ANCHOR, LANDMARK = 0, 1, 2, 3, 4
NAMES = {P2P: "point_to_point", PLANE: "point_to_plane", SMOOTH: "smoothness", ANCHOR: "anchor",
         LANDMARK: "landmark"}


def blocks(mod, pts, state, mask=None, weights=None):
    return mod.energy_blocks(*energy_args(pts, state), weights or all_weights(), mask)


def test_all_five_block_kinds_appear(host_impl, oracle):
    pts, _ = sphere_patch()
    r = blocks(host_impl, pts, graph_state(oracle, pts, 13, 0))
    assert set(np.asarray(r["blocks"]["kind"]).tolist()) == {P2P, PLANE, SMOOTH, ANCHOR, LANDMARK}


def test_energies_are_the_squared_rows(host_impl, oracle):
    pts, _ = sphere_patch()
    r = blocks(host_impl, pts, graph_state(oracle, pts, 13, 1))
    b = r["blocks"]
    values, rows, kind = np.asarray(b["values"]), np.asarray(b["rows"]), np.asarray(b["kind"])
    sq = np.array([np.sum(values[i, :rows[i]] ** 2) for i in range(len(rows))])
    for k, name in NAMES.items():
        assert r["energies"][name] == pytest.approx(sq[kind == k].sum(), rel=1e-12)
    assert r["energies"]["total"] == pytest.approx(sq.sum(), rel=1e-12)


def test_a_masked_vertex_leaves_the_data_term_only(host_impl, oracle):
    pts, _ = sphere_patch()
    state = graph_state(oracle, pts, 13, 2)
    full = blocks(host_impl, pts, state)
    mask = np.ones(len(pts), dtype=np.int64)
    mask[[6, 7, 8, 20]] = 0
    part = blocks(host_impl, pts, state, mask=mask)
    data = lambda r: int(np.isin(np.asarray(r["blocks"]["kind"]), [P2P, PLANE]).sum())  # noqa: E731
    assert data(full) - data(part) == 2 * 4
    for name in ("point_to_point", "point_to_plane"):
        assert part["energies"][name] < full["energies"][name]
    for name in ("smoothness", "anchor", "landmark"):
        assert part["energies"][name] == pytest.approx(full["energies"][name], rel=1e-14)


def test_jacobians_match_central_differences(host_impl, oracle):
    pts, _ = sphere_patch()
    n = 13
    state = graph_state(oracle, pts, n, 3)
    rng = np.random.default_rng(5)
    direction = rng.normal(size=6 * n)
    direction /= np.linalg.norm(direction)
    eps = 1e-6

    def at(step):
        moved = oracle.retract_deformation(state["rotations"], state["translations"], state["nodes"], step)
        s = dict(state, rotations=moved["rotations"], translations=moved["translations"])
        return blocks(host_impl, pts, s)["blocks"]

    base, plus, minus = blocks(host_impl, pts, state)["blocks"], at(eps * direction), at(-eps * direction)
    rows, slots, nodes = np.asarray(base["rows"]), np.asarray(base["slots"]), np.asarray(base["nodes"])
    jac = np.asarray(base["jacobians"])
    worst = 0.0
    for i in range(len(rows)):
        r = rows[i]
        predicted = sum(jac[i, s, :r, :] @ direction[6 * nodes[i, s]:6 * nodes[i, s] + 6] for s in range(slots[i]))
        measured = (np.asarray(plus["values"])[i, :r] - np.asarray(minus["values"])[i, :r]) / (2 * eps)
        if np.abs(measured).max() == 0.0:
            continue
        np.testing.assert_allclose(measured, predicted, rtol=1e-4, atol=1e-6)
        worst = max(worst, np.abs(measured - predicted).max() / max(np.abs(measured).max(), np.abs(predicted).max()))
    assert worst < 1e-4


def test_robust_kernels_follow_their_closed_forms(impl):
    s = np.linspace(0.0, 9.0, 61)
    t = 0.4
    huber = np.asarray(impl.robust_weights("huber", s, t))
    gm = np.asarray(impl.robust_weights("geman_mcclure", s, t))
    for w in (huber, gm):
        assert w[0] == 1.0 and np.all(np.diff(w) <= 0.0) and np.all((w > 0) & (w <= 1))
    inside = s <= t * t
    np.testing.assert_allclose(huber[inside], 1.0, rtol=0, atol=0)
    np.testing.assert_allclose(huber[~inside], t / np.sqrt(s[~inside]), rtol=1e-15, atol=0)
    np.testing.assert_allclose(gm, 1.0 / (1.0 + s / (t * t)) ** 2, rtol=1e-15, atol=0)
    assert np.all(gm[~inside] < huber[~inside])
    np.testing.assert_allclose(np.asarray(impl.robust_weights("none", s, t)), 1.0, rtol=0, atol=0)
