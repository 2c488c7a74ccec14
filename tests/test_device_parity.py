"""The device path against the FP64 oracle on identical inputs.

Integer and index results (seeds, bound nodes, edges, triangle ids, rejection
codes) must be identical; FP64 results that the device computes with the same
operation order (normals, sampling distances, closest points) must be
bit-identical; reductions and solves are held to stated tolerances. The one
documented exception: when two triangles are at exactly the same distance
from a query, the device picks the lower triangle index and the oracle the
first one its hierarchy visits. On a mesh without duplicated vertices the
tied triangles share the tied feature, so points and normals agree to the last
bit or two; on the reference's UV sphere, whose poles are 2 x segments
coincident but distinct vertices with different normals, a tie can pick a
different pole copy and so a different normal. Even on the welded sphere the
mirror symmetry leaves exact ties between triangles with different closest
points; the route-level comparisons therefore run on a jittered welded sphere,
where exact ties have probability zero.
"""

from __future__ import annotations

import functools
import json

import numpy as np
import pytest

from meshes import all_weights, bulge, energy_args, graph_state, route_settings, sphere_patch, uv_sphere, welded_uv_sphere

pytestmark = pytest.mark.gpu


def _noisy_sphere(seed=0, rings=16, segments=24):
    points, faces = welded_uv_sphere(rings=rings, segments=segments)
    rng = np.random.default_rng(seed)
    return points + rng.normal(scale=0.01, size=points.shape), faces


def test_vertex_normals_are_bit_identical(device, oracle):
    points, faces = _noisy_sphere(1)
    assert np.array_equal(device.vertex_normals(points, faces), oracle.vertex_normals(points, faces))


def test_closest_points_match_exactly_except_equidistant_ties(device, oracle):
    points, faces = _noisy_sphere(2)
    rng = np.random.default_rng(3)
    queries = rng.normal(size=(2000, 3))
    queries *= rng.uniform(0.3, 1.7, size=(2000, 1)) / np.linalg.norm(queries, axis=1, keepdims=True)
    queries = np.vstack([queries, points[:200]])  # exact vertex hits: ties by construction
    d = device.closest_points(points, faces, queries)
    o = oracle.closest_points(points, faces, queries)
    assert np.array_equal(d["distances"], o["distances"])
    same = d["triangles"] == o["triangles"]
    assert np.array_equal(d["points"][same], o["points"][same])
    assert np.array_equal(d["normals"][same], o["normals"][same])
    # Where the triangle differs the two are exact distance ties on a shared
    # feature: same point and normal up to the rounding of the two corner
    # orders.
    assert np.array_equal(d["distances"][~same], o["distances"][~same])
    assert np.allclose(d["points"][~same], o["points"][~same], rtol=0, atol=1e-15)
    assert np.allclose(d["normals"][~same], o["normals"][~same], rtol=0, atol=1e-15)
    assert np.array_equal(d["vertex_normals"], o["vertex_normals"])


def _morton10(q):
    """Interleaved bits of three 10-bit cell coordinates (x highest), as the
    device's hierarchy keys use them."""
    code = 0
    for bit in range(9, -1, -1):
        for axis in range(3):
            code = (code << 1) | ((int(q[axis]) >> bit) & 1)
    return code


def _deepest_internal_node(keys):
    """Deepest internal node of the binary radix tree over sorted unique keys
    (each node splits its range at the highest bit in which it differs)."""
    keys = [int(k) for k in keys]

    def walk(lo, hi, depth):
        if hi == lo:
            return -1
        bit = (keys[lo] ^ keys[hi]).bit_length() - 1
        split = next(i for i in range(lo, hi + 1) if (keys[i] >> bit) & 1)
        return max(depth, walk(lo, split - 1, depth + 1), walk(split, hi, depth + 1))

    return walk(0, len(keys) - 1, 0)


def _deep_hierarchy_mesh(seed=11):
    """Separate small triangles whose centroids put the hierarchy's keys on a
    chain that splits one Morton bit per level (30 levels), above 4,097 keys
    sharing one cell (about 12 more levels of index bits): a tree deeper than
    a query stack of 96 entries covers by the 7 W - 6 bound."""
    rng = np.random.default_rng(seed)
    centroids = [rng.uniform(0.05, 0.95, size=3) for _ in range(4096)]  # cell 0
    centroids.append(np.zeros(3))  # sets the low corner of the centroid box
    for level in range(10):
        for axis in range(3):
            c = np.full(3, 0.5)
            c[axis] = 2.0**level + 0.5
            centroids.append(c)
    centroids.append(np.full(3, 1024.0))  # sets the high corner: cells are unit cubes
    centroids = np.array(centroids)
    # A small triangle around each centroid, in a random orientation.
    points, faces = [], []
    for t, c in enumerate(centroids):
        u, v = rng.normal(size=3), rng.normal(size=3)
        a, b = 0.01 * u, 0.01 * v
        points += [c + a, c + b, c - a - b]
        faces.append([3 * t, 3 * t + 1, 3 * t + 2])
    points, faces = np.array(points), np.array(faces)
    # The keys as the device builds them: cells of the centroids' box, then
    # the triangle index in the low bits.
    cent = points[faces].mean(axis=1)
    lo, hi = cent.min(0), cent.max(0)
    cells = np.clip(np.floor((cent - lo) / (hi - lo) * 1024.0), 0, 1023).astype(np.int64)
    index_bits = max(1, int(len(faces) - 1).bit_length())
    keys = sorted((_morton10(q) << index_bits) | t for t, q in enumerate(cells))
    return points, faces, _deepest_internal_node(keys)


def test_closest_points_on_a_deep_hierarchy(device, oracle):
    points, faces, depth = _deep_hierarchy_mesh()
    wide_levels = depth // 3 + 1
    assert 7 * wide_levels - 6 > 96  # the query runs with its 128-entry stack
    rng = np.random.default_rng(12)
    queries = np.vstack([rng.uniform(-0.5, 1.5, size=(3000, 3)),        # at the deep cluster
                         rng.uniform(-1.0, 520.0, size=(1000, 3))])     # along the chain
    d = device.closest_points(points, faces, queries)
    o = oracle.closest_points(points, faces, queries)
    assert np.array_equal(d["triangles"], o["triangles"])
    assert np.array_equal(d["distances"], o["distances"])
    assert np.array_equal(d["points"], o["points"])


def test_farthest_point_seeds_are_identical(device, oracle):
    rng = np.random.default_rng(4)
    points = rng.normal(size=(5000, 3))
    for count, first in ((1, 0), (25, 0), (400, 17)):
        assert np.array_equal(device.farthest_point_seeds(points, count, first),
                              oracle.farthest_point_seeds(points, count, first))


@pytest.mark.parametrize("n", [70_000, 150_000, 300_000])
def test_farthest_point_seeds_are_identical_at_every_block_size(device, oracle, n):
    # The sampler stores the points in blocks of 64 x ceil(n / 65536) (in
    # rounds of 4 above 4 per lane): these sizes take 2, 3 and 8 per lane.
    rng = np.random.default_rng(n)
    points = rng.normal(size=(n, 3)) * np.array([1.0, 0.3, 2.0])
    assert np.array_equal(device.farthest_point_seeds(points, 300, 5), oracle.farthest_point_seeds(points, 300, 5))


def test_farthest_point_seeds_break_exact_ties_like_the_oracle(device, oracle):
    # A lattice has many equal squared distances; the repeated points make
    # the running distances all zero before the last seed, after which every
    # pick is the lowest index.
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(12), indexing="ij"), -1).reshape(-1, 3)
    points = np.concatenate([g, g[::7]]).astype(np.float64)
    for count in (50, len(points)):
        assert np.array_equal(device.farthest_point_seeds(points, count, 3),
                              oracle.farthest_point_seeds(points, count, 3))


def test_bindings_and_edges_are_identical(device, oracle):
    rng = np.random.default_rng(5)
    points = rng.normal(size=(3000, 3))
    nodes = points[oracle.farthest_point_seeds(points, 120)]
    for k in (1, 2, 4):
        d = device.bind_vertices(points, nodes, k)
        o = oracle.bind_vertices(points, nodes, k)
        for key in ("nodes", "weights", "counts"):
            assert np.array_equal(d[key], o[key]), (k, key)
    for k in (4, 8):
        assert np.array_equal(device.build_node_edges(nodes, k), oracle.build_node_edges(nodes, k))


def _awkward_node_sets():
    rng = np.random.default_rng(11)
    lattice = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij"), -1).reshape(-1, 3)
    yield "lattice (exact ties)", lattice.astype(np.float64), rng.uniform(-1, 6, size=(800, 3))
    line = np.zeros((40, 3))
    line[:, 0] = np.linspace(0.0, 1.0, 40)
    yield "collinear", line, rng.normal(size=(500, 3))
    yield "coincident", np.ones((7, 3)), rng.normal(size=(300, 3))
    yield "far queries", rng.normal(size=(200, 3)), rng.normal(size=(400, 3)) * 50.0
    yield "three nodes", rng.normal(size=(3, 3)), rng.normal(size=(100, 3))
    sheet = rng.uniform(size=(5000, 3)) * np.array([10.0, 10.0, 0.01])
    yield "thin sheet", sheet[oracle_free_fps(sheet, 700)], sheet
    # Above 2^24 query-node pairs the device searches the grid one wave per
    # query instead of every pair: the same awkward shapes at that size.
    big = np.stack(np.meshgrid(*(np.arange(17),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    yield "large lattice (exact ties)", big, rng.uniform(-1, 17, size=(4000, 3))
    wide = rng.uniform(size=(20000, 3)) * np.array([30.0, 30.0, 0.01])
    yield "large thin sheet", wide[oracle_free_fps(wide, 4500)], wide[:5000]


@functools.lru_cache(maxsize=None)
def _awkward_cases():
    return list(_awkward_node_sets())


def oracle_free_fps(points, count):
    # Greedy farthest points (numpy), only to spread the nodes of a fixture.
    chosen = [0]
    d = np.sum((points - points[0]) ** 2, axis=1)
    for _ in range(count - 1):
        j = int(np.argmax(d))
        chosen.append(j)
        d = np.minimum(d, np.sum((points - points[j]) ** 2, axis=1))
    return np.asarray(chosen)


@pytest.mark.parametrize("case", range(8))
def test_grid_searches_match_brute_force_on_awkward_node_sets(device, oracle, case):
    # The device searches nodes through a uniform grid (the kNN edges and the
    # level transfer: every pair up to 2^24 pairs, a wave per query beyond);
    # the oracle scans them all. Lattices give exact distance ties, coincident
    # nodes a degenerate box, far queries long searches, a thin sheet flat
    # cells.
    name, nodes, queries = _awkward_cases()[case]
    for k in (1, 4):
        d = device.bind_vertices(queries, nodes, k)
        o = oracle.bind_vertices(queries, nodes, k)
        for key in ("nodes", "weights", "counts"):
            assert np.array_equal(d[key], o[key]), (name, k, key)
    for k in (4, 8, 16):
        assert np.array_equal(device.build_node_edges(nodes, k), oracle.build_node_edges(nodes, k)), (name, k)
    rng = np.random.default_rng(case)
    angles = rng.normal(size=(len(nodes), 3)) * 0.3
    rotations = np.stack([_rotation(a) for a in angles])
    translations = rng.normal(size=(len(nodes), 3)) * 0.1
    td = device.transfer_level(nodes, rotations, translations, queries)
    to = oracle.transfer_level(nodes, rotations, translations, queries)
    # The blend's arithmetic is not bit for bit the oracle's (as in the level
    # transfer test above); a different set of nearest nodes would show as an
    # O(1) difference.
    for key in ("rotations", "translations"):
        assert np.allclose(td[key], to[key], rtol=0, atol=1e-12), (name, key)


def _rotation(w):
    t = np.linalg.norm(w)
    if t == 0.0:
        return np.eye(3)
    k = w / t
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K


@pytest.mark.parametrize("neighbours", [4, 8])
def test_template_graph_is_identical(device, oracle, neighbours):
    points, faces = uv_sphere(rings=24, segments=20)
    d = device.build_template_graph(points, faces, 3, 40, 2, 4, True, neighbours)
    o = oracle.build_template_graph(points, faces, 3, 40, 2, 4, False, neighbours)
    assert d["normalization"]["scale"] == o["normalization"]["scale"]
    for left, right in zip(d["levels"], o["levels"]):
        assert np.array_equal(left["node_template_indices"], right["node_template_indices"])
        assert np.array_equal(left["node_positions"], right["node_positions"])
        for key in ("nodes", "weights", "counts"):
            assert np.array_equal(left["bindings"][key], right["bindings"][key])
        assert np.array_equal(left["edges"], right["edges"])


def test_deformation_operators_match(device, oracle):
    points, _ = sphere_patch()
    state = graph_state(oracle, points, 13, seed=6)
    b = state["binding"]
    args = (points, b["nodes"], b["weights"], b["counts"], state["rotations"], state["translations"], state["nodes"])
    assert np.allclose(device.evaluate_deformation(*args), oracle.evaluate_deformation(*args), rtol=0, atol=1e-15)
    delta = np.random.default_rng(7).normal(scale=0.05, size=13 * 6)
    rd = device.retract_deformation(state["rotations"], state["translations"], state["nodes"], delta)
    ro = oracle.retract_deformation(state["rotations"], state["translations"], state["nodes"], delta)
    assert np.allclose(rd["rotations"], ro["rotations"], rtol=0, atol=1e-15)
    assert np.array_equal(rd["translations"], ro["translations"])
    fine = points[oracle.farthest_point_seeds(points, 30)]
    td = device.transfer_level(state["nodes"], state["rotations"], state["translations"], fine)
    to = oracle.transfer_level(state["nodes"], state["rotations"], state["translations"], fine)
    assert np.allclose(td["rotations"], to["rotations"], rtol=0, atol=1e-14)
    assert np.allclose(td["translations"], to["translations"], rtol=0, atol=1e-14)


@pytest.mark.parametrize("masked", [False, True])
def test_fused_system_matches_the_oracle(device, oracle, masked):
    points, _ = sphere_patch()
    state = graph_state(oracle, points, 13, seed=8)
    mask = None
    if masked:
        mask = np.ones(len(points), dtype=np.int64)
        mask[[2, 5, 6]] = 0
    accepted = np.ones(len(points), dtype=np.int64)
    accepted[[9, 10]] = 0
    robust = np.random.default_rng(9).uniform(0.2, 1.0, size=len(points))
    args = energy_args(points, state, accepted, robust)
    d = device.fused_system(*args, all_weights(), mask)
    o = oracle.fused_system(*args, all_weights(), mask)
    assert np.array_equal(d["edges"], o["edges"])
    scale = np.abs(o["diagonal"]).max()
    assert np.allclose(d["diagonal"], o["diagonal"], rtol=0, atol=1e-12 * scale)
    assert np.allclose(d["edge_blocks"], o["edge_blocks"], rtol=0, atol=1e-12 * scale)
    assert np.allclose(d["rhs"], o["rhs"], rtol=0, atol=1e-12 * max(np.abs(o["rhs"]).max(), 1.0))
    for key, value in o["energies"].items():
        assert d["energies"][key] == pytest.approx(value, rel=1e-12, abs=1e-300), key


@pytest.mark.parametrize("precision", ["mixed", "float"])
def test_fp32_fused_systems_match_the_oracle_to_fp32_accuracy(device, oracle, precision):
    """The mixed form's FP32 blocks are the FP64 ones rounded (half an FP32
    ulp of each entry); the all-FP32 form computes and accumulates them in
    FP32 (a few ulps of the largest entry). The right-hand side and the
    energies stay FP64 in both."""
    points, _ = sphere_patch()
    state = graph_state(oracle, points, 13, seed=8)
    robust = np.random.default_rng(9).uniform(0.2, 1.0, size=len(points))
    args = energy_args(points, state, None, robust)
    d = device.fused_system(*args, all_weights(), None, precision)
    o = oracle.fused_system(*args, all_weights(), None)
    assert np.array_equal(d["edges"], o["edges"])
    scale = np.abs(o["diagonal"]).max()
    if precision == "mixed":
        f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)  # noqa: E731
        assert np.array_equal(d["diagonal"], f32(o["diagonal"]))
        assert np.array_equal(d["edge_blocks"], f32(o["edge_blocks"]))
    else:
        assert np.allclose(d["diagonal"], o["diagonal"], rtol=0, atol=1e-6 * scale)
        assert np.allclose(d["edge_blocks"], o["edge_blocks"], rtol=0, atol=1e-6 * scale)
        assert not np.array_equal(d["diagonal"], o["diagonal"])
    assert np.allclose(d["rhs"], o["rhs"], rtol=0, atol=1e-12 * max(np.abs(o["rhs"]).max(), 1.0))
    for key, value in o["energies"].items():
        assert d["energies"][key] == pytest.approx(value, rel=1e-12, abs=1e-300), key


def test_solve_matches_the_direct_solve(device, oracle):
    points, _ = sphere_patch()
    state = graph_state(oracle, points, 13, seed=10)
    system = oracle.fused_system(*energy_args(points, state), all_weights())
    for lam, scaling in ((1e-3, True), (1e-6, False), (1.0, True)):
        o = oracle.solve_normal_equations(system["diagonal"], system["edges"], system["edge_blocks"], system["rhs"],
                                          lam, scaling)
        d = device.solve_normal_equations(system["diagonal"], system["edges"], system["edge_blocks"], system["rhs"],
                                          lam, scaling)
        assert d["ok"] and o["ok"]
        assert np.allclose(d["delta"], o["delta"], rtol=0, atol=1e-9 * max(1.0, np.abs(o["delta"]).max()))
        m = device.solve_normal_equations(system["diagonal"], system["edges"], system["edge_blocks"],
                                          system["rhs"], lam, scaling, precision="mixed")
        assert m["ok"]
        assert np.allclose(m["delta"], o["delta"], rtol=0, atol=1e-8 * max(1.0, np.abs(o["delta"]).max()))


def test_correspondence_filters_match(device, oracle):
    points, faces = _noisy_sphere(11)
    rng = np.random.default_rng(12)
    queries = points * rng.uniform(0.9, 1.1, size=(len(points), 1))
    normals = oracle.vertex_normals(queries, faces)
    normals[::7] *= -1.0
    mask = (rng.random(len(points)) > 0.05).astype(np.int64)
    kwargs = dict(distance_threshold=0.05, min_normal_cosine=0.3, min_slope=0.2, mask=mask)
    d = device.find_correspondences(points, faces, queries, normals, **kwargs)
    o = oracle.find_correspondences(points, faces, queries, normals, **kwargs)
    assert np.array_equal(d["rejection"], o["rejection"])
    for key in ("total", "accepted", "rejected_distance", "rejected_normal", "rejected_slope", "rejected_masked"):
        assert d["counts"][key] == o["counts"][key], key
    assert d["counts"]["distance_p95"] == o["counts"]["distance_p95"]
    assert d["counts"]["distance_max"] == o["counts"]["distance_max"]
    assert d["counts"]["distance_mean"] == pytest.approx(o["counts"]["distance_mean"], rel=1e-13)


def _jittered_pair(seed=0, rings=18, segments=24, amplitude=0.09):
    template, faces = welded_uv_sphere(rings=rings, segments=segments)
    rng = np.random.default_rng(seed)
    template = template + rng.normal(scale=2e-3, size=template.shape)
    target = bulge(template, amplitude) + rng.normal(scale=2e-3, size=template.shape)
    return template, faces, target


def _rms_max(a, b, diag):
    err = np.linalg.norm(np.asarray(a) - np.asarray(b), axis=1) / diag
    return float(np.sqrt(np.mean(err ** 2))), float(err.max())


@pytest.mark.parametrize("levels,nodes", [(1, 60), (2, 80)])
def test_route_a_matches_the_oracle(device, oracle, levels, nodes):
    template, faces, target = _jittered_pair(seed=levels)
    kwargs = dict(graph_levels=levels, nodes_initial=nodes, growth=2)
    d = device.solve_route_a(template, faces, target, faces, route_settings(levels), **kwargs)
    o = oracle.solve_route_a(template, faces, target, faces, route_settings(levels), **kwargs)
    diag = float(np.linalg.norm(template.max(0) - template.min(0)))
    rms, worst = _rms_max(d["vertices"], o["vertices"], diag)
    assert rms <= 1e-5 and worst <= 1e-4, (rms, worst)
    rd, ro = json.loads(d["report"]), json.loads(o["report"])
    assert [l["inner_iterations"] for l in rd["levels"]] == [l["inner_iterations"] for l in ro["levels"]]
    assert [l["outer_iterations"] for l in rd["levels"]] == [l["outer_iterations"] for l in ro["levels"]]
    assert np.allclose(d["trace_energy"], o["trace_energy"], rtol=1e-9)
    assert rd["route"] == "route_a" and rd["precision"] == "double"
    assert np.array_equal(d["level_nodes"], o["level_nodes"])


def test_route_b_mixed_precision_matches_the_oracle(device, oracle):
    template, faces, target = _jittered_pair(seed=7)
    kwargs = dict(graph_levels=2, nodes_initial=80, growth=2)
    d = device.solve_route_b(template, faces, target, faces, route_settings(2), pcg_tolerance=1e-12,
                             pcg_max_iterations=5000, precision="mixed", **kwargs)
    o = oracle.solve_route_a(template, faces, target, faces, route_settings(2), **kwargs)
    diag = float(np.linalg.norm(template.max(0) - template.min(0)))
    rms, worst = _rms_max(d["vertices"], o["vertices"], diag)
    assert rms <= 1e-5 and worst <= 1e-4, (rms, worst)
    assert json.loads(d["report"])["precision"] == "float"


@pytest.mark.parametrize("precision", ["double", "mixed", "float"])
def test_route_c_matches_the_oracle(device, oracle, precision):
    template, faces, target = _jittered_pair(seed=11)
    kwargs = dict(graph_levels=2, nodes_initial=80, growth=2)
    d = device.solve_route_c(template, faces, target, faces, route_settings(2), pcg_tolerance=1e-12,
                             pcg_max_iterations=5000, precision=precision, **kwargs)
    o = oracle.solve_route_c(template, faces, target, faces, route_settings(2), **kwargs)
    diag = float(np.linalg.norm(template.max(0) - template.min(0)))
    rms, worst = _rms_max(d["vertices"], o["vertices"], diag)
    assert rms <= 1e-5 and worst <= 1e-4, (rms, worst)
    rd, ro = json.loads(d["report"]), json.loads(o["report"])
    assert rd["route"] == "route_c"
    assert [l["inner_iterations"] for l in rd["levels"]] == [l["inner_iterations"] for l in ro["levels"]]
