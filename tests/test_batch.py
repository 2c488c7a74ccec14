"""The lockstep batch (nricp_dp.Batch): uneven jobs solved together through
shared launches -- one launch per stage of a damping attempt for every job,
one persistent PCG over all jobs' block rows, the accept/reject decisions on
the device, each level's ticks replayed from graphs.

The contract checked here: a job's result is the same bits alone and in the
batch, with and without graph capture; with the PCG tile of the single-job
session (one node slot per lane group while every level has 129 to 2,048
nodes) it is the same bits as the single-job session's; the routes match the
FP64 oracle within the parity bound; and a batch refuses what it cannot run.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

from nricp_amd import workloads

KWARGS = dict(graph_levels=2, growth=2, graph_neighbours=8, pcg_tolerance=1e-12, pcg_max_iterations=2000)


def _jobs():
    # Uneven sizes (1,002 to 2,562 template vertices), every level above 128
    # nodes, so the single-job session runs its multi-group PCG with one node
    # slot per lane group (the packed PCG's tile at pcg_slots=1).
    specs = [(10, 14, 280), (16, 22, 400), (12, 16, 300), (14, 20, 360), (12, 20, 264)]
    return [workloads.sphere_job(f, g, seed=300 + k, nodes_final=n, graph_levels=2, name=f"batch{k}")
            for k, (f, g, n) in enumerate(specs)]


def _args(job, outer=3):
    settings = workloads.level_settings(job.graph_levels, outer=outer, inner=1)
    return (job.template_points, job.template_faces, job.target_points, job.target_faces, settings)


def _batch(device, jobs, form, slots=1, graphs=True, route="route_b"):
    batch = device.Batch(pcg_slots=slots, use_graphs=graphs)
    for job in jobs:
        batch.add(*_args(job), nodes_initial=job.nodes_initial, route=route, precision=form, **KWARGS)
    results = batch.run()
    stats = batch.stats()
    batch.close()
    return results, stats


def _same(a, b, label):
    assert np.array_equal(a["vertices"], b["vertices"]), label
    ra, rb = json.loads(a["report"]), json.loads(b["report"])
    assert ra["levels"] == rb["levels"], label
    assert ra["energy"] == rb["energy"], label
    assert ra["correspondence"] == rb["correspondence"], label
    assert np.array_equal(a["level_nodes"], b["level_nodes"]), label


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["float", "mixed", "double"])
def test_batch_matches_the_single_job_session_bit_for_bit(device, form):
    jobs = _jobs()
    print(f"batch vs session: {len(jobs)} jobs, form {form}", flush=True)
    packed, stats = _batch(device, jobs, form, slots=1)
    assert stats["graphs_built"] >= 2 and stats["graph_replays"] == stats["ticks"] > 0
    for job, got in zip(jobs, packed):
        alone = device.solve_route_b(*_args(job), nodes_initial=job.nodes_initial, precision=form, **KWARGS)
        _same(alone, got, (job.name, form))


@pytest.mark.gpu
@pytest.mark.parametrize("slots", [1, 2])
def test_a_job_gives_the_same_bits_alone_and_in_the_batch(device, slots):
    jobs = _jobs()
    packed, _ = _batch(device, jobs, "float", slots=slots)
    for k in (0, 3):
        alone, _ = _batch(device, [jobs[k]], "float", slots=slots)
        _same(alone[0], packed[k], (jobs[k].name, slots))


@pytest.mark.gpu
def test_graph_replay_gives_the_same_bits_as_direct_launches(device):
    jobs = _jobs()[:3]
    replayed, rs = _batch(device, jobs, "mixed", slots=2, graphs=True)
    direct, ds = _batch(device, jobs, "mixed", slots=2, graphs=False)
    assert rs["graph_replays"] == rs["ticks"] and ds["graph_replays"] == 0 and ds["ticks"] == rs["ticks"]
    for job, a, b in zip(jobs, replayed, direct):
        _same(a, b, job.name)


@pytest.mark.gpu
@pytest.mark.parametrize("route,form", [("route_a", "double"), ("route_b", "float"), ("route_b", "mixed")])
def test_batch_routes_match_the_oracle(device, oracle, route, form):
    # Jittered spheres: no exact closest-point ties, so the device and the
    # oracle take the same accept/reject path; coarse levels of 40 nodes
    # exercise the one-group solve inside the packed launch.
    jobs = [workloads.sphere_job(f, g, seed=400 + k, nodes_final=n, graph_levels=2, name=f"oracle{k}")
            for k, (f, g, n) in enumerate([(8, 12, 80), (12, 16, 120), (10, 16, 100)])]
    packed, _ = _batch(device, jobs, form, slots=2, route=route)
    for job, got in zip(jobs, packed):
        ref = oracle.solve_route_a(*_args(job), nodes_initial=job.nodes_initial, **KWARGS)
        diag = float(np.linalg.norm(job.template_points.max(0) - job.template_points.min(0)))
        err = np.linalg.norm(got["vertices"] - ref["vertices"], axis=1) / diag
        rms = float(np.sqrt(np.mean(err ** 2)))
        print(f"{job.name} {route}/{form}: rms {rms:.3e} max {err.max():.3e}", flush=True)
        assert rms <= 1e-5 and err.max() <= 1e-4, (job.name, rms, err.max())
        rg, rr = json.loads(got["report"]), json.loads(ref["report"])
        assert [l["outer_iterations"] for l in rg["levels"]] == [l["outer_iterations"] for l in rr["levels"]]
        assert [l["inner_iterations"] for l in rg["levels"]] == [l["inner_iterations"] for l in rr["levels"]]
        assert rg["correspondence"]["accepted"] == rr["correspondence"]["accepted"]
        assert np.array_equal(got["level_nodes"], ref["level_nodes"])


@pytest.mark.gpu
def test_batch_refuses_what_it_cannot_run(device):
    job = _jobs()[0]
    b = device.Batch()
    with pytest.raises(ValueError):
        b.add(*_args(job), nodes_initial=job.nodes_initial, route="route_c", **KWARGS)
    b.add(*_args(job), nodes_initial=job.nodes_initial, precision="float", **KWARGS)
    b.add(*_args(job), nodes_initial=job.nodes_initial, precision="double", **KWARGS)
    with pytest.raises(ValueError):
        b.run()  # one precision per batch
    odd = device.Batch(pcg_slots=3)
    odd.add(*_args(job), nodes_initial=job.nodes_initial, **KWARGS)
    with pytest.raises(ValueError):
        odd.run()  # the PCG tile takes one or two node slots per lane group
