"""The run report's JSON: every field survives a round trip, and non-finite
numbers become null (the reference's tests/test_report.py, restated; the
byte-exact layout is pinned by tests/test_golden.py)."""

from __future__ import annotations

import json
import math

ENERGY = {"point_to_point": 0.75, "point_to_plane": 1.5, "smoothness": 0.0625, "anchor": 0.0, "landmark": 2.5}
CORR = {"total": 5000, "accepted": 4970, "rejected_distance": 20, "rejected_normal": 6, "rejected_slope": 4,
        "rejected_masked": 0, "distance_mean": 0.021, "distance_p95": 0.08, "distance_max": 0.31}
LEVELS = [
    {"level": 0, "nodes": 125, "outer_iterations": 8, "inner_iterations": 5, "energy_start": 9.0,
     "energy_end": 2.75, "correspondence_mean": 0.6, "converged": True},
    {"level": 1, "nodes": 250, "outer_iterations": 10, "inner_iterations": 7, "energy_start": 2.75,
     "energy_end": 0.875, "correspondence_mean": 0.2, "converged": False},
]


def test_every_field_round_trips(host_impl):
    settings = {"levels": 2.0, "level0_smoothness": 0.05}
    r = json.loads(host_impl.serialize_report("route_b", "float", 0.5, 2.25, ENERGY, CORR, LEVELS, settings))
    assert (r["route"], r["precision"], r["version"]) == ("route_b", "float", host_impl.__version__)
    assert (r["setup_seconds"], r["solve_seconds"]) == (0.5, 2.25)
    assert r["energy"] == dict(ENERGY, total=sum(ENERGY.values()))
    assert all(r["correspondence"][k] == v for k, v in CORR.items())
    assert r["levels"] == LEVELS
    assert r["settings"] == settings


def test_non_finite_numbers_are_null(host_impl):
    energy = dict(ENERGY, point_to_point=math.nan, landmark=math.inf)
    corr = {k: 0 if isinstance(v, int) else 0.0 for k, v in CORR.items()}
    r = json.loads(host_impl.serialize_report("route_a", "double", 0.0, 0.0, energy, corr, [], {}))
    assert r["energy"]["point_to_point"] is None and r["energy"]["landmark"] is None
    assert r["energy"]["total"] is None
    assert r["levels"] == [] and r["settings"] == {}
