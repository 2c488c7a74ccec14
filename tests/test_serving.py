"""The request loop (nricp_dp.Server): jobs of different sizes admitted by
slot and device-memory budget, running on the device at the same time.

The contract checked here: a job's result is the same bits whether it runs
alone (solve_route_b) or among others in the server; admission never lets
the reserved memory exceed the budget; every admitted job's session fits in
what admission reserved for it; bad jobs fail at submission.
"""

from __future__ import annotations

import json
import threading
import time

import numpy as np
import pytest

from nricp_amd import workloads

KWARGS = dict(graph_levels=2, growth=2, graph_neighbours=8, pcg_tolerance=1e-12, pcg_max_iterations=2000)


def _jobs():
    # Uneven sizes: templates of 642 to 2,562 vertices, targets denser.
    specs = [(8, 12, 40), (16, 22, 100), (10, 14, 60), (12, 20, 60), (8, 16, 30), (14, 18, 80)]
    return [workloads.sphere_job(f, g, seed=100 + k, nodes_final=n, graph_levels=2, name=f"serve{k}")
            for k, (f, g, n) in enumerate(specs)]


def _args(job):
    settings = workloads.level_settings(job.graph_levels, outer=3, inner=1)
    return (job.template_points, job.template_faces, job.target_points, job.target_faces, settings)


def _alone(device, job):
    return device.solve_route_b(*_args(job), nodes_initial=job.nodes_initial, precision="mixed", **KWARGS)


@pytest.mark.gpu
def test_results_are_the_same_alone_and_in_the_mix(device):
    jobs = _jobs()
    alone = [_alone(device, j) for j in jobs]
    with device.Server(max_active=3) as server:
        tickets = [server.submit(*_args(j), nodes_initial=j.nodes_initial, precision="mixed", **KWARGS) for j in jobs]
        mixed = [server.result(t) for t in tickets]
        assert server.stats()["finished"] == len(jobs)
    for job, a, m in zip(jobs, alone, mixed):
        assert np.array_equal(a["vertices"], m["vertices"]), job.name
        ra, rm = json.loads(a["report"]), json.loads(m["report"])
        assert ra["energy"] == rm["energy"], job.name
        assert ra["levels"] == rm["levels"], job.name
        assert ra["correspondence"] == rm["correspondence"], job.name
        assert 0 < m["device_bytes"] <= m["estimated_bytes"], (job.name, m["device_bytes"], m["estimated_bytes"])
        assert m["run_seconds"] > 0.0 and m["queued_seconds"] >= 0.0


@pytest.mark.gpu
def test_every_route_and_form_gives_the_same_bits_in_the_mix(device):
    # Route A (FP64 to 1e-12), route B in both forms and route C (local and
    # global steps) running side by side.
    jobs = _jobs()[:4]
    specs = [("route_a", "double"), ("route_b", "double"), ("route_c", "mixed"), ("route_c", "double")]
    alone = []
    for job, (route, form) in zip(jobs, specs):
        solve = getattr(device, f"solve_{route}")
        extra = {} if route == "route_a" else {"precision": form}
        alone.append(solve(*_args(job), nodes_initial=job.nodes_initial, **extra, **KWARGS))
    with device.Server(max_active=4) as server:
        tickets = [server.submit(*_args(j), nodes_initial=j.nodes_initial, route=route, precision=form, **KWARGS)
                   for j, (route, form) in zip(jobs, specs)]
        mixed = [server.result(t) for t in tickets]
    for job, spec, a, m in zip(jobs, specs, alone, mixed):
        assert np.array_equal(a["vertices"], m["vertices"]), (job.name, spec)
        assert json.loads(a["report"])["route"] == json.loads(m["report"])["route"] == spec[0]


@pytest.mark.gpu
def test_admission_keeps_reserved_memory_within_the_budget(device):
    jobs = _jobs()
    estimates = []
    with device.Server(max_active=4) as probe:
        for j in jobs[:1]:
            t = probe.submit(*_args(j), nodes_initial=j.nodes_initial, **KWARGS)
            estimates.append(probe.result(t)["estimated_bytes"])
    # Room for about two of the smallest jobs at once.
    budget = 2.5 * estimates[0]
    with device.Server(max_active=4, device_budget_bytes=budget) as server:
        tickets = [server.submit(*_args(j), nodes_initial=j.nodes_initial, **KWARGS) for j in jobs]
        seen = []
        stop = threading.Event()

        def watch():
            while not stop.is_set():
                seen.append(server.stats())
                time.sleep(0.002)

        watcher = threading.Thread(target=watch)
        watcher.start()
        try:
            results = [server.result(t) for t in tickets]
        finally:
            stop.set()
            watcher.join()
        for s in seen:
            assert s["reserved_bytes"] <= s["budget_bytes"] + 1e-6
            assert s["active"] <= 4
        assert all(r["device_bytes"] <= r["estimated_bytes"] for r in results)
    for job, r in zip(jobs, results):
        assert np.array_equal(r["vertices"], _alone(device, job)["vertices"]), job.name


@pytest.mark.gpu
def test_bad_jobs_fail_at_submission(device):
    job = _jobs()[0]
    with device.Server(max_active=1, device_budget_bytes=1e6) as server:
        with pytest.raises(ValueError):  # larger than the whole budget
            server.submit(*_args(job), nodes_initial=job.nodes_initial, **KWARGS)
    with device.Server(max_active=1) as server:
        faces = job.template_faces.copy()
        faces[0, 0] = len(job.template_points) + 5
        with pytest.raises(ValueError):
            server.submit(job.template_points, faces, job.target_points, job.target_faces,
                          workloads.level_settings(2), nodes_initial=job.nodes_initial, **KWARGS)
        with pytest.raises(ValueError):
            server.result(12345)


def test_server_is_part_of_the_module_surface(host_module):
    assert hasattr(host_module, "Server")
    doc = host_module.Server.submit.__doc__
    for name in ("template_points", "level_settings", "graph_neighbours", "route", "precision"):
        assert name in doc
