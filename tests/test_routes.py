"""The routes end to end on synthetic deformations (the reference's
tests/test_route_a.py, test_route_b.py, test_route_c.py and test_api_smoke.py,
restated). Route A and route C run on the FP64 oracle here and on the device
under ``-m gpu``; route B and the step-by-step session exist on the device
only."""

from __future__ import annotations

import json
import re
import sys
from pathlib import Path

import numpy as np
import pytest

from meshes import bulge, route_settings, uv_sphere

GOLDEN = Path(__file__).resolve().parent / "golden"


def run(mod, name, template, faces, target, levels=1, target_faces=None, nodes_initial=60, overrides=None,
        **kwargs):
    return getattr(mod, name)(template, faces, target, faces if target_faces is None else target_faces,
                              route_settings(levels, **(overrides or {})), graph_levels=levels,
                              nodes_initial=nodes_initial, growth=2, **kwargs)


def mean_error(vertices, truth):
    return float(np.linalg.norm(np.asarray(vertices) - truth, axis=1).mean())


@pytest.fixture(params=["solve_route_a", "solve_route_c"])
def route(request):
    return request.param


# ------------------------------------------------------------ every route --

def test_one_level_fit_reaches_the_surface_and_reports_it(impl, route):
    t, f = uv_sphere(rings=18, segments=24)
    truth = bulge(t)
    r = run(impl, route, t, f, truth)
    rep = json.loads(r["report"])
    before, after = mean_error(t, truth), mean_error(r["vertices"], truth)
    assert after < 0.75 * before, (before, after)
    assert rep["correspondence"]["distance_mean"] < 0.1 * before
    lv = rep["levels"][0]
    assert lv["nodes"] == 60 and lv["outer_iterations"] >= 1 and lv["inner_iterations"] >= 1
    assert lv["energy_end"] < lv["energy_start"]
    assert rep["energy"]["total"] > 0.0 and rep["correspondence"]["accepted"] > 0
    assert rep["route"] == route.replace("solve_", "") and rep["version"] == impl.__version__


def test_three_levels_are_not_worse_than_one(impl, route):
    t, f = uv_sphere(rings=18, segments=24)
    truth = bulge(t, amplitude=0.12)
    one = run(impl, route, t, f, truth, levels=1, nodes_initial=240)
    three = run(impl, route, t, f, truth, levels=3, nodes_initial=60)
    assert mean_error(three["vertices"], truth) <= 1.1 * mean_error(one["vertices"], truth)
    assert np.asarray(three["level_nodes"]).tolist() == [60, 120, 240]
    assert [l["nodes"] for l in json.loads(three["report"])["levels"]] == [60, 120, 240]


def test_a_loose_outer_tolerance_stops_the_level_early(impl, route):
    t, f = uv_sphere(rings=16, segments=20)
    r = run(impl, route, t, f, bulge(t, amplitude=0.06), nodes_initial=80,
            overrides={"max_outer_iterations": 8, "outer_tolerance": 0.5})
    lv = json.loads(r["report"])["levels"][0]
    assert lv["converged"] is True and lv["outer_iterations"] < 8 and lv["energy_end"] < lv["energy_start"]


def test_heavy_landmarks_pull_their_vertices_home(impl, route):
    t, f = uv_sphere(rings=16, segments=20)
    truth = bulge(t, amplitude=0.10)
    idx = np.array([12, 85, 150, 233, 301], dtype=np.int64)
    out = truth[idx] - t.mean(0)
    goal = truth[idx] + 0.02 * out / np.linalg.norm(out, axis=1, keepdims=True)
    common = dict(levels=2, nodes_initial=80, landmark_indices=idx, landmark_targets=goal)
    weak = run(impl, route, t, f, truth, overrides={"landmark": 1.0}, **common)
    strong = run(impl, route, t, f, truth, overrides={"landmark": 2000.0}, **common)
    residual = lambda r: np.linalg.norm(np.asarray(r["vertices"])[idx] - goal, axis=1).mean()  # noqa: E731
    assert residual(strong) < residual(weak)
    if route == "solve_route_a":
        assert residual(strong) < 1e-2 * np.linalg.norm(t.max(0) - t.min(0))


def test_a_tiny_robust_scale_suppresses_the_data_term(impl, route):
    t, f = uv_sphere(rings=16, segments=20)
    truth = bulge(t, amplitude=0.10)
    plain = run(impl, route, t, f, truth, nodes_initial=80)
    damped = run(impl, route, t, f, truth, nodes_initial=80,
                 overrides={"robust_kernel": "geman_mcclure", "robust_threshold": 1e-6})
    motion = lambda r: np.linalg.norm(np.asarray(r["vertices"]) - t, axis=1).mean()  # noqa: E731
    assert motion(damped) < 0.5 * motion(plain)


def floating_cap(template, faces, lift=0.35, height=0.85):
    clean = bulge(template, amplitude=0.10)
    cap = np.where(clean[:, 2] > height)[0]
    on_cap = np.isin(faces, cap).any(axis=1)
    moved = np.unique(faces[on_cap])
    remap = -np.ones(len(clean), dtype=np.int64)
    remap[moved] = len(clean) + np.arange(len(moved))
    pts = np.vstack([clean, clean[moved] + np.array([0.0, 0.0, lift])])
    return clean, cap, pts, np.vstack([faces[~on_cap], remap[faces[on_cap]]])


def test_the_distance_threshold_drops_a_detached_patch(impl, route):
    t, f = uv_sphere(rings=18, segments=24)
    clean, cap, pts, tf = floating_cap(t, f)
    common = dict(levels=2, nodes_initial=80, target_faces=tf)
    loose = run(impl, route, t, f, pts, **common)
    tight = run(impl, route, t, f, pts, distance_threshold=0.1, **common)
    err = lambda r: np.linalg.norm(np.asarray(r["vertices"])[cap] - clean[cap], axis=1).mean()  # noqa: E731
    assert err(tight) < 0.5 * err(loose)
    assert json.loads(tight["report"])["correspondence"]["rejected_distance"] > 0


def test_masked_vertices_follow_the_regulariser(impl, route):
    t, f = uv_sphere(rings=16, segments=20)
    truth = bulge(t, amplitude=0.10)
    mask = np.ones(len(t), dtype=np.int64)
    mask[:40] = 0
    r = run(impl, route, t, f, truth, nodes_initial=80, mask=mask)
    assert json.loads(r["report"])["correspondence"]["rejected_masked"] == 40
    fitted = np.asarray(r["vertices"])
    assert np.linalg.norm(fitted[:40] - t[:40], axis=1).mean() < 1.5 * np.linalg.norm(truth - t, axis=1).mean()


# --------------------------------------------------------------- route C --

def rotations(r):
    return np.asarray(r["rotations"]).reshape(-1, 3, 3)


def test_route_c_sits_still_when_there_is_nothing_to_fit(impl):
    t, f = uv_sphere(rings=12, segments=16)
    r = run(impl, "solve_route_c", t, f, t, levels=2)
    assert np.abs(rotations(r) - np.eye(3)).max() < 1e-4
    assert np.abs(np.asarray(r["translations"])).max() < 1e-4
    assert np.abs(np.asarray(r["vertices"]) - t).max() < 1e-4


def test_route_c_rotations_are_proper(impl):
    t, f = uv_sphere(rings=12, segments=16)
    R = rotations(run(impl, "solve_route_c", t, f, bulge(t, 0.08), levels=2))
    assert np.isfinite(R).all()
    assert np.abs(np.einsum("nij,nkj->nik", R, R) - np.eye(3)).max() < 1e-4
    d = np.linalg.det(R)
    assert d.min() > 0.999 and d.max() < 1.001


def test_route_c_carries_a_translation_with_its_translations(impl):
    base, f = uv_sphere(rings=14, segments=18)
    t = base * np.array([1.4, 1.0, 0.75])
    shift = np.array([0.0, 0.05, 0.0])
    r = run(impl, "solve_route_c", t, f, t + shift, levels=2, nodes_initial=80)
    fit = np.asarray(r["vertices"])
    moved = fit.mean(0) - t.mean(0)
    assert np.linalg.norm(moved) > 0.15 * np.linalg.norm(shift)
    assert moved @ shift / np.linalg.norm(moved) / np.linalg.norm(shift) > 0.99
    ratio = np.sort(np.linalg.eigvalsh(np.cov((fit - fit.mean(0)).T))) / \
        np.sort(np.linalg.eigvalsh(np.cov((t - t.mean(0)).T)))
    assert np.all(ratio > 0.95) and np.all(ratio < 1.02)
    R = rotations(r)
    angles = np.degrees(np.arccos(np.clip((np.trace(R, axis1=1, axis2=2) - 1) / 2, -1, 1)))
    assert np.median(angles) < 6.0


def test_route_c_lands_near_route_a(impl):
    t, f = uv_sphere(rings=14, segments=18)
    truth = bulge(t, 0.09)
    c = mean_error(run(impl, "solve_route_c", t, f, truth, levels=2, nodes_initial=80)["vertices"], truth)
    a = mean_error(run(impl, "solve_route_a", t, f, truth, levels=2, nodes_initial=80)["vertices"], truth)
    assert c < 0.75 * mean_error(t, truth) and c < 1.10 * a


def test_route_c_levels_descend_and_name_the_route(impl):
    t, f = uv_sphere(rings=12, segments=16)
    r = run(impl, "solve_route_c", t, f, bulge(t, 0.08), levels=2)
    rep = json.loads(r["report"])
    assert rep["route"] == "route_c" and rep["version"] == impl.__version__
    assert [l["nodes"] for l in rep["levels"]] == list(np.asarray(r["level_nodes"]))
    for lv in rep["levels"]:
        assert lv["inner_iterations"] > 0 and lv["energy_end"] <= lv["energy_start"]


# ------------------------------------------------- device routes, session --

@pytest.mark.gpu
def test_the_device_is_reported(device):
    info = device.device_info()
    assert set(info) == {"available", "device"} and info["available"] and "gfx950" in info["device"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "double"])
def test_route_b_agrees_with_route_a(device, oracle, precision):
    t, f = uv_sphere(rings=14, segments=18)
    truth = bulge(t, 0.09)
    b = run(device, "solve_route_b", t, f, truth, levels=2, nodes_initial=80, precision=precision,
            pcg_tolerance=1e-10, pcg_max_iterations=2000)
    a = run(oracle, "solve_route_a", t, f, truth, levels=2, nodes_initial=80)
    e0, eb, ea = mean_error(t, truth), mean_error(b["vertices"], truth), mean_error(a["vertices"], truth)
    assert np.isfinite(np.asarray(b["vertices"])).all()
    assert eb < 0.75 * e0 and ea < 0.75 * e0
    assert abs(eb - ea) < 0.05 * ea


@pytest.mark.gpu
def test_route_b_with_the_reference_defaults_still_fits(device):
    """The reference's route B defaults (tolerance 1e-5, 60 iterations)."""
    t, f = uv_sphere(rings=12, segments=16)
    truth = bulge(t, 0.08)
    r = run(device, "solve_route_b", t, f, truth, levels=2)
    assert mean_error(r["vertices"], truth) < 0.75 * mean_error(t, truth)
    rep = json.loads(r["report"])
    assert rep["route"] == "route_b" and rep["precision"] == "float"
    for lv in rep["levels"]:
        assert lv["energy_end"] <= lv["energy_start"]


@pytest.mark.gpu
def test_route_b_repeats_bit_for_bit(device):
    """Every device reduction has a fixed order, so two runs agree exactly."""
    t, f = uv_sphere(rings=12, segments=16)
    truth = bulge(t, 0.08)
    one = run(device, "solve_route_b", t, f, truth, levels=2)
    two = run(device, "solve_route_b", t, f, truth, levels=2)
    assert np.array_equal(np.asarray(one["vertices"]), np.asarray(two["vertices"]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["route_b", "route_c", "route_a"])
def test_the_session_steps_to_the_batch_result(device, name):
    t, f = uv_sphere(rings=10, segments=14)
    truth = bulge(t, 0.07)
    settings = route_settings(2)
    args = dict(graph_levels=2, nodes_initial=50, growth=2)
    s = device.RouteBSession(t, f, truth, f, settings, route=name, **args)
    steps, early = 0, None
    while not s.finished():
        p = s.step()
        steps += 1
        assert p["level"] < 2 and steps < 5000
        if steps == 3:
            part = np.asarray(s.vertices())
            assert part.shape == t.shape and np.isfinite(part).all()
            early = mean_error(part, truth)
    stepped = np.asarray(s.result()["vertices"])
    batch = np.asarray(getattr(device, "solve_" + name)(t, f, truth, f, settings, **args)["vertices"])
    assert np.array_equal(stepped, batch)
    assert early is not None and early > mean_error(stepped, truth) and mean_error(stepped, truth) < mean_error(t, truth)
    assert json.loads(s.report())["route"] == name


@pytest.mark.gpu
def test_the_session_refuses_an_unknown_route(device):
    t, f = uv_sphere(rings=6, segments=8)
    with pytest.raises(ValueError):
        device.RouteBSession(t, f, t, f, route_settings(1), graph_levels=1, nodes_initial=20, route="route_q")


# ------------------------------------------------------------ module info --

def test_version_and_build_info(host_module):
    assert re.fullmatch(r"\d+\.\d+\.\d+", host_module.__version__)
    info = host_module.build_info()
    for key in ("version", "build_type", "cxx_compiler", "cuda_toolkit", "cuda_architectures",
                "accelerated_precision", "python_version", "pybind11_version"):
        assert isinstance(info[key], str) and info[key], key
    assert info["version"] == host_module.__version__
    assert info["python_version"].startswith("%d.%d" % sys.version_info[:2])
    assert "gfx950" in info["cuda_architectures"] and info["cuda_toolkit"].startswith("HIP")


@pytest.mark.gpu
def test_the_device_graph_builder_is_the_oracles_bit_for_bit(device, oracle):
    for rings, segments, nodes, growth in ((12, 16, 40, 2), (24, 20, 100, 2), (32, 24, 150, 3)):
        t, f = uv_sphere(rings=rings, segments=segments)
        h = oracle.build_template_graph(t, f, 3, nodes, growth)
        d = device.build_template_graph(t, f, 3, nodes, growth)
        assert h["normalization"]["scale"] == d["normalization"]["scale"]
        for left, right in zip(h["levels"], d["levels"]):
            for key in ("node_template_indices", "node_positions", "edges"):
                assert np.array_equal(np.asarray(left[key]), np.asarray(right[key])), key
            for key in ("nodes", "weights", "counts"):
                assert np.array_equal(np.asarray(left["bindings"][key]), np.asarray(right["bindings"][key])), key


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["route_a", "double", "mixed", "float"])
def test_a_level_above_one_persistent_solve_matches_the_oracle(device, form):
    """A level of 16,500 nodes is more than one persistent PCG launch holds
    on this device (16,384), so its solves run as the multi-launch PCG. One
    Gauss-Newton step of a sphere onto a bumped sphere against the oracle's
    route A (its result is the fixture tests/golden/large_level_route_a.npz,
    written by tests/golden/make_large_level.py: the oracle takes a minute at
    this size): vertices within 1e-5 RMS and 1e-4 max of the box diagonal,
    the same level size, correspondence counts and outer iterations. The
    same job submitted to a server gives the same bits."""
    sys.path.insert(0, str(GOLDEN))
    import make_large_level as large

    fixture = np.load(GOLDEN / "large_level_route_a.npz")
    job = large.job()
    args, kwargs = large.call_args(job)
    if form == "route_a":
        d = device.solve_route_a(*args, **kwargs)
    else:
        kwargs.update(precision=form, pcg_tolerance=1e-10, pcg_max_iterations=5000)
        d = device.solve_route_b(*args, **kwargs)
    expected = job.template_points + fixture["displacement"].astype(np.float64)
    diag = float(np.linalg.norm(job.template_points.max(0) - job.template_points.min(0)))
    e = np.linalg.norm(np.asarray(d["vertices"]) - expected, axis=1) / diag
    rms, worst = float(np.sqrt(np.mean(e ** 2))), float(e.max())
    print(f"16,500-node level, {form}: rms {rms:.3e} max {worst:.3e} of the diagonal")
    assert rms <= 1e-5 and worst <= 1e-4, (rms, worst)
    assert list(d["level_nodes"]) == list(fixture["level_nodes"]) == [large.NODES]
    report = json.loads(d["report"])
    assert [report["correspondence"][k] for k in large.COUNTS] == list(fixture["correspondence"])
    assert [l["outer_iterations"] for l in report["levels"]] == list(fixture["outer_iterations"])
    if form == "double":
        with device.Server(max_active=1) as server:
            served = server.result(server.submit(*args, route="route_b", **kwargs))
        assert np.array_equal(served["vertices"], d["vertices"])


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["route_a", "mixed", "float"])
def test_speculative_attempts_change_nothing(device, form):
    """Each damping attempt's successor solved beside it (levels of 1,024
    nodes or more) gives the same bits, the same report and the same solve
    counts as the attempts one after another: two levels of 2,000 and 4,000
    nodes, five correspondence updates each."""
    from nricp_amd import workloads

    job = workloads.sphere_job(41, 60, 5, nodes_final=4_000, graph_levels=2)
    settings = workloads.level_settings(2, outer=5, inner=1)
    args = (job.template_points, job.template_faces, job.target_points, job.target_faces, settings)
    # A small initial damping: the first steps overshoot and are rejected.
    kwargs = dict(graph_levels=2, nodes_initial=2_000, growth=2, graph_neighbours=8, initial_damping=1e-7)
    if form == "route_a":
        kwargs.update(route="route_a")
    else:
        kwargs.update(precision=form, pcg_tolerance=1e-8, pcg_max_iterations=50)

    def run(speculative):
        session = device.RouteBSession(*args, speculative_attempts=speculative, **kwargs)
        session.run()
        return session.result(), session.counters()

    (on, on_counts), (off, off_counts) = run(True), run(False)
    assert np.array_equal(on["vertices"], off["vertices"])
    assert json.loads(on["report"])["levels"] == json.loads(off["report"])["levels"]
    for key in ("pcg_launches", "pcg_iterations", "gauss_newton_solves"):
        assert on_counts[key] == off_counts[key], key
    # Rejected attempts are what the speculation hides: this job has some.
    outer = sum(l["outer_iterations"] for l in json.loads(on["report"])["levels"])
    assert on_counts["gauss_newton_solves"] > outer
