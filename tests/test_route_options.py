"""The device routes against the oracle with the reference's route options
switched on: IRLS weighting with either kernel (src/solver/route_a.cpp:19-38,
src/energy/robust.cpp:30-58), landmarks, a vertex mask, a finite distance
threshold and the normal / slope filters -- the options the reference's own
route tests exercise (tests/test_route_a.py:134-280). Each case runs route A
(FP64) and route B (mixed) against the oracle's route A, and route C against
the oracle's route C, on jittered inputs (no exact closest-point ties), and
requires the vertices within 1e-5 (RMS) and 1e-4 (max) of the box diagonal,
the same inner and outer iteration counts at every level and the same final
correspondence counts."""

from __future__ import annotations

import json

import numpy as np
import pytest

from meshes import bulge, route_settings, welded_uv_sphere


def _pair(seed):
    template, faces = welded_uv_sphere(rings=18, segments=24)
    rng = np.random.default_rng(seed)
    template = template + rng.normal(scale=2e-3, size=template.shape)
    target = bulge(template, 0.09) + rng.normal(scale=2e-3, size=template.shape)
    # The top cap lifted outward (the triangles around it stretch): its
    # correspondences are far, so a finite distance threshold rejects them
    # and IRLS down-weights them.
    cap = np.where(target[:, 2] > 0.8)[0]
    target[cap] += np.array([0.0, 0.0, 0.12])
    return template, faces, target


def _landmarks(template, target):
    idx = np.array([5, 77, 160, 251, 333], dtype=np.int64)
    return idx, target[idx] + 0.01


def _mask(n, seed):
    rng = np.random.default_rng(seed)
    mask = (rng.random(n) > 0.1).astype(np.int64)
    return mask


CASES = {
    "huber": (dict(robust_kernel="huber", robust_threshold=0.02), {}),
    "geman_mcclure": (dict(robust_kernel="geman_mcclure", robust_threshold=0.03), {}),
    "landmarks": ({}, {"landmarks": True}),
    "mask": ({}, {"mask": True}),
    "distance_threshold": ({}, {"distance_threshold": 0.03}),
    "everything": (dict(robust_kernel="geman_mcclure", robust_threshold=0.05),
                   {"landmarks": True, "mask": True, "distance_threshold": 0.035, "min_normal_cosine": 0.2,
                    "min_slope": 0.05}),
}


# The local/global route with correspondences rejected by distance is
# sensitive to its inputs on most seeds here: the oracle's own final vertices
# move by up to 4e-2 of the diagonal under a 1e-13 perturbation of the target
# (route A moves by 1e-13). No second implementation can track such a
# trajectory, so route C's threshold case uses a threshold and seed the oracle
# shows well-posed (test_oracle_cases_are_well_posed checks every case used).
ROUTE_C_OPTIONS = {"distance_threshold": {"distance_threshold": 0.035}}
SEEDS = {"solve_route_a": 31, "solve_route_b": 32, "solve_route_c": 33}


def _run(mod, name, case, seed, target_noise=0.0, **extra):
    overrides, options = CASES[case]
    if name == "solve_route_c" and case in ROUTE_C_OPTIONS:
        options = dict(options, **ROUTE_C_OPTIONS[case])
    template, faces, target = _pair(seed)
    if target_noise > 0.0:
        target = target + np.random.default_rng(99).normal(scale=target_noise, size=target.shape)
    kwargs = dict(graph_levels=2, nodes_initial=60, growth=2)
    if options.get("landmarks"):
        idx, goal = _landmarks(template, target)
        kwargs.update(landmark_indices=idx, landmark_targets=goal)
    if options.get("mask"):
        kwargs["mask"] = _mask(len(template), seed)
    for key in ("distance_threshold", "min_normal_cosine", "min_slope"):
        if key in options:
            kwargs[key] = options[key]
    kwargs.update(extra)
    result = getattr(mod, name)(template, faces, target, faces, route_settings(2, **overrides), **kwargs)
    return result, template


def _compare(d, o, template):
    diag = float(np.linalg.norm(template.max(0) - template.min(0)))
    err = np.linalg.norm(np.asarray(d["vertices"]) - np.asarray(o["vertices"]), axis=1) / diag
    rms, worst = float(np.sqrt(np.mean(err ** 2))), float(err.max())
    assert rms <= 1e-5 and worst <= 1e-4, (rms, worst)
    rd, ro = json.loads(d["report"]), json.loads(o["report"])
    assert [l["outer_iterations"] for l in rd["levels"]] == [l["outer_iterations"] for l in ro["levels"]]
    assert [l["inner_iterations"] for l in rd["levels"]] == [l["inner_iterations"] for l in ro["levels"]]
    for key in ("accepted", "rejected_distance", "rejected_normal", "rejected_slope", "rejected_masked"):
        assert rd["correspondence"][key] == ro["correspondence"][key], key
    return rd, ro


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_options_change_the_fit(oracle, case):
    """Each option makes a difference on this input (else the parity check
    below would not test it)."""
    plain, template = _run(oracle, "solve_route_a", "mask", 31, mask=None)
    with_option, _ = _run(oracle, "solve_route_a", case, 31)
    assert np.abs(np.asarray(plain["vertices"]) - np.asarray(with_option["vertices"])).max() > 1e-4
    report = json.loads(with_option["report"])["correspondence"]
    if CASES[case][1].get("distance_threshold"):
        assert report["rejected_distance"] > 0
    if CASES[case][1].get("mask"):
        assert report["rejected_masked"] > 0


@pytest.mark.parametrize("route", ["solve_route_a", "solve_route_c"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_cases_are_well_posed(oracle, route, case):
    """A 1e-13 perturbation of the target moves the oracle's result by at
    most 1e-10 of the diagonal: the trajectory the device is held to is a
    property of the input, not of rounding."""
    seeds = [SEEDS[route]] + ([SEEDS["solve_route_b"]] if route == "solve_route_a" else [])
    for seed in seeds:
        a, template = _run(oracle, route, case, seed)
        b, _ = _run(oracle, route, case, seed, target_noise=1e-13)
        diag = float(np.linalg.norm(template.max(0) - template.min(0)))
        moved = np.linalg.norm(np.asarray(a["vertices"]) - np.asarray(b["vertices"]), axis=1).max() / diag
        assert moved < 1e-10, (seed, moved)


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_device_route_a_matches_the_oracle_with_options(device, oracle, case):
    d, template = _run(device, "solve_route_a", case, 31)
    o, _ = _run(oracle, "solve_route_a", case, 31)
    _compare(d, o, template)
    assert np.allclose(d["trace_energy"], o["trace_energy"], rtol=1e-9)


# The mixed form iterates in FP32 but refines against the FP64 system, so a
# solve to 1e-12 is the FP64 solve and the trajectory is the oracle's.
@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_device_route_b_mixed_matches_the_oracle_with_options(device, oracle, case):
    d, template = _run(device, "solve_route_b", case, 32, pcg_tolerance=1e-12, pcg_max_iterations=5000,
                       precision="mixed")
    o, _ = _run(oracle, "solve_route_a", case, 32)
    _compare(d, o, template)
    assert np.allclose(d["trace_energy"], o["trace_energy"], rtol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_device_route_c_matches_the_oracle_with_options(device, oracle, case):
    d, template = _run(device, "solve_route_c", case, 33, pcg_tolerance=1e-12, pcg_max_iterations=5000,
                       precision="double")
    o, _ = _run(oracle, "solve_route_c", case, 33)
    _compare(d, o, template)
