"""Template-side stages checked against independent expectations -- closed
forms, a numpy restatement, or an invariant the stage must keep -- in the
manner of the reference's tests/test_core_geometry.py. Device-backed entry
points run through ``impl`` (the oracle here, the device under ``-m gpu``);
host-only ones through ``host_impl`` (the oracle and the device module's host
code)."""

from __future__ import annotations

import numpy as np
import pytest

from meshes import rotation_about, uv_sphere


def numpy_fps(points, count, first=0):
    chosen = [first]
    nearest = np.sum((points - points[first]) ** 2, axis=1)
    for _ in range(1, count):
        nxt = int(np.argmax(nearest))
        chosen.append(nxt)
        nearest = np.minimum(nearest, np.sum((points - points[nxt]) ** 2, axis=1))
    return np.asarray(chosen)


# ------------------------------------------------------------ normalisation --

def test_normalization_centres_the_box_and_scales_its_diagonal_to_one(host_impl):
    rng = np.random.default_rng(11)
    cloud = rng.normal(size=(400, 3)) * np.array([2.0, 0.5, 1.5]) - 3.0
    z = host_impl.normalization(cloud)
    centre = np.asarray(z["center"]).reshape(3)
    np.testing.assert_allclose(centre, 0.5 * (cloud.min(0) + cloud.max(0)), rtol=0, atol=1e-14)
    unit = (cloud - centre) * z["scale"]
    assert np.linalg.norm(unit.max(0) - unit.min(0)) == pytest.approx(1.0, abs=1e-12)
    np.testing.assert_allclose(unit / z["scale"] + centre, cloud, rtol=0, atol=1e-12)


def test_validation_names_bad_indices_and_non_finite_points(host_impl):
    tri = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    ok = host_impl.validate_mesh(tri, np.array([[0, 1, 2]]))
    assert ok["ok"] and list(ok["problems"]) == []
    bad = host_impl.validate_mesh(tri, np.array([[2, 1, 5]]))
    assert not bad["ok"] and any("out of range" in p for p in bad["problems"])
    nan = tri.copy()
    nan[2, 0] = np.inf
    bad = host_impl.validate_mesh(nan, np.array([[0, 1, 2]]))
    assert not bad["ok"] and any("not finite" in p for p in bad["problems"])


# ----------------------------------------------------------------- cleaning --

def test_cleaning_counts_duplicates_degenerates_and_orphans(host_impl):
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [3, 0, 0], [9, 9, 9]], dtype=np.float64)
    faces = np.array([[0, 1, 2], [1, 3, 2], [2, 1, 0], [1, 1, 2], [0, 1, 4]], dtype=np.int64)
    r = host_impl.clean_mesh(pts, faces)
    assert (r["duplicate_faces"], r["degenerate_faces"], r["isolated_vertices"]) == (1, 2, 2)
    assert np.asarray(r["faces"]).shape == (2, 3) and np.asarray(r["points"]).shape == (4, 3)
    assert np.asarray(r["faces"]).max() < 4
    assert list(np.asarray(r["faces"])[0]) == [0, 1, 2]  # the first winding survives


def test_faces_sharing_only_an_edge_are_not_duplicates(host_impl):
    tet = np.array([[0.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    faces = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]], dtype=np.int64)
    r = host_impl.clean_mesh(tet, faces)
    assert (r["duplicate_faces"], r["degenerate_faces"]) == (0, 0)
    r = host_impl.clean_mesh(tet, np.vstack([faces, faces[:, [2, 1, 0]]]))
    assert r["duplicate_faces"] == 4 and np.asarray(r["faces"]).shape == (4, 3)


# ----------------------------------------------------------------- sampling --

def test_farthest_point_seeds_follow_the_plain_algorithm(impl):
    cloud = np.random.default_rng(12).normal(size=(900, 3))
    seeds = np.asarray(impl.farthest_point_seeds(cloud, 30))
    assert np.array_equal(seeds, numpy_fps(cloud, 30))
    assert np.array_equal(seeds, np.asarray(impl.farthest_point_seeds(cloud, 30)))
    assert len(np.asarray(impl.farthest_point_seeds(cloud, 10_000))) == len(cloud)


def test_farthest_point_seeds_start_where_asked(impl):
    cloud = np.random.default_rng(13).normal(size=(150, 3))
    seeds = np.asarray(impl.farthest_point_seeds(cloud, 12, 9))
    assert seeds[0] == 9 and np.array_equal(seeds, numpy_fps(cloud, 12, 9))


# ------------------------------------------------------------------ binding --

def test_binding_weights_are_the_normalised_squared_falloff(impl):
    nodes = np.array([[0.5, 0, 0], [0, 1.5, 0], [0, 0, 2.5], [3.5, 0, 0]])
    r = impl.bind_vertices(np.zeros((1, 3)), nodes)
    w = np.asarray(r["weights"])[0]
    f = (1.0 - np.array([0.5, 1.5, 2.5, 3.5]) / 3.5) ** 2
    np.testing.assert_allclose(w, f / f.sum(), rtol=0, atol=1e-15)
    assert w[3] == 0.0 and np.asarray(r["nodes"])[0, 0] == 0


def test_binding_sums_to_one_and_its_heaviest_node_is_the_nearest(impl):
    rng = np.random.default_rng(14)
    pts, nodes = rng.normal(size=(250, 3)), rng.normal(size=(15, 3))
    r = impl.bind_vertices(pts, nodes)
    w, n = np.asarray(r["weights"]), np.asarray(r["nodes"])
    assert (np.asarray(r["counts"]) == 4).all()
    np.testing.assert_allclose(w.sum(1), 1.0, rtol=0, atol=1e-14)
    nearest = np.argmin(((pts[:, None] - nodes[None]) ** 2).sum(-1), axis=1)
    assert np.array_equal(n[np.arange(len(pts)), w.argmax(1)], nearest)
    again = impl.bind_vertices(pts, nodes)
    assert np.array_equal(np.asarray(again["nodes"]), n) and np.array_equal(np.asarray(again["weights"]), w)


def test_binding_with_fewer_nodes_than_slots(impl):
    r = impl.bind_vertices(np.array([[0.0, 0, 0], [2.0, 2, 2]]), np.array([[0.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0]]))
    assert (np.asarray(r["counts"]) == 3).all()
    np.testing.assert_allclose(np.asarray(r["weights"]).sum(1), 1.0, rtol=0, atol=1e-14)


# ---------------------------------------------------------------- the graph --

def test_template_graph_levels_are_fps_prefixes_in_normalised_units(impl):
    pts, faces = uv_sphere(rings=16, segments=24)
    g = impl.build_template_graph(pts, faces, levels=3, nodes_initial=40, growth=2)
    sizes = [len(np.asarray(l["node_template_indices"])) for l in g["levels"]]
    assert sizes == [40, 80, 160]
    centre = np.asarray(g["normalization"]["center"]).reshape(3)
    finest = np.asarray(g["levels"][-1]["node_template_indices"])
    for level in g["levels"]:
        idx = np.asarray(level["node_template_indices"])
        assert np.array_equal(idx, finest[: len(idx)])
        assert len(np.unique(idx)) == len(idx)
        np.testing.assert_allclose(np.asarray(level["node_positions"]), (pts[idx] - centre) * g["normalization"]["scale"],
                                   rtol=0, atol=1e-12)
        assert np.asarray(level["bindings"]["nodes"]).max() < len(idx)


# -------------------------------------------------------------- deformation --

def _rigid(angle, axis, shift):
    return rotation_about(np.asarray(axis, dtype=float), angle), np.asarray(shift, dtype=float)


def test_identity_and_rigid_fields_evaluate_exactly(impl):
    pts, _ = uv_sphere(rings=9, segments=11)
    nodes = pts[::11]
    b = impl.bind_vertices(pts, nodes)
    eye = np.repeat(np.eye(3)[None], len(nodes), 0)
    out = impl.evaluate_deformation(pts, b["nodes"], b["weights"], b["counts"], eye, np.zeros((len(nodes), 3)), nodes)
    np.testing.assert_allclose(out, pts, rtol=0, atol=1e-12)
    R, t = _rigid(0.55, [0.2, 1.0, 0.3], [0.1, -0.4, 0.25])
    out = impl.evaluate_deformation(pts, b["nodes"], b["weights"], b["counts"], np.repeat(R[None], len(nodes), 0),
                                    nodes @ R.T + t - nodes, nodes)
    np.testing.assert_allclose(out, pts @ R.T + t, rtol=0, atol=1e-12)


def test_level_transfer_keeps_a_rigid_field(impl):
    pts, faces = uv_sphere(rings=10, segments=14)
    g = impl.build_template_graph(pts, faces, levels=2, nodes_initial=30, growth=3)
    coarse = np.asarray(g["levels"][0]["node_positions"])
    fine = np.asarray(g["levels"][1]["node_positions"])
    R, t = _rigid(0.3, [1.0, 0.2, -0.1], [0.05, 0.3, -0.2])
    r = impl.transfer_level(coarse, np.repeat(R[None], len(coarse), 0), coarse @ R.T + t - coarse, fine)
    np.testing.assert_allclose(np.asarray(r["rotations"]), np.repeat(R[None], len(fine), 0), rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.asarray(r["translations"]), fine @ R.T + t - fine, rtol=0, atol=1e-12)


def test_level_transfer_does_not_collapse_opposite_half_turns(impl):
    pts, faces = uv_sphere(rings=10, segments=14)
    g = impl.build_template_graph(pts, faces, levels=2, nodes_initial=30, growth=3)
    coarse = np.asarray(g["levels"][0]["node_positions"])
    R = rotation_about(np.array([0.0, 0.0, 1.0]), np.deg2rad(172.0))
    rots = np.stack([R if i % 2 == 0 else R.T for i in range(len(coarse))])
    r = impl.transfer_level(coarse, rots, np.zeros((len(coarse), 3)), np.asarray(g["levels"][1]["node_positions"]))
    got = np.asarray(r["rotations"])
    angle = np.arccos(np.clip((np.trace(got, axis1=1, axis2=2) - 1) / 2, -1, 1))
    assert np.isfinite(got).all() and (angle > np.deg2rad(90.0)).all()


def test_rigid_fit_recovers_a_transform_and_copes_with_a_line(host_impl):
    rng = np.random.default_rng(15)
    src = rng.normal(size=(25, 3))
    R, t = _rigid(0.8, [0.3, -1.0, 0.4], [0.3, 0.1, -0.5])
    fit = host_impl.fit_rigid(src, src @ R.T + t, np.ones(len(src)))
    np.testing.assert_allclose(np.asarray(fit["rotation"]).reshape(3, 3), R, rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.asarray(fit["translation"]).reshape(3), t, rtol=0, atol=1e-12)
    line = np.stack([np.linspace(-1, 1, 7), np.zeros(7), np.zeros(7)], 1)
    fit = host_impl.fit_rigid(line, line @ R.T + t, np.ones(7))
    Rf, tf = np.asarray(fit["rotation"]).reshape(3, 3), np.asarray(fit["translation"]).reshape(3)
    assert np.isfinite(Rf).all() and np.isfinite(tf).all()
    assert np.abs(line @ Rf.T + tf - (line @ R.T + t)).max() < 1e-9


def test_retraction_and_tangent_difference_round_trip(impl, host_impl):
    pts, _ = uv_sphere(rings=8, segments=12)
    nodes = pts[::13]
    n = len(nodes)
    eye, zero = np.repeat(np.eye(3)[None], n, 0), np.zeros((n, 3))
    same = impl.retract_deformation(eye, zero, nodes, np.zeros(6 * n))
    assert np.array_equal(np.asarray(same["rotations"]), eye) and np.array_equal(np.asarray(same["translations"]), zero)
    step = np.random.default_rng(16).normal(scale=1e-3, size=6 * n)
    moved = impl.retract_deformation(eye, zero, nodes, step)
    back = host_impl.tangent_difference(eye, zero, moved["rotations"], moved["translations"])
    np.testing.assert_allclose(np.asarray(back), step, rtol=0, atol=1e-12)


# ----------------------------------------------------------- closest points --

def test_closest_points_on_a_fine_sphere_are_radial(impl):
    pts, faces = uv_sphere(rings=32, segments=64)
    rng = np.random.default_rng(17)
    d = rng.normal(size=(250, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    radius = rng.uniform(0.45, 1.55, size=(250, 1))
    r = impl.closest_points(pts, faces, d * radius)
    np.testing.assert_allclose(np.asarray(r["distances"]), np.abs(radius[:, 0] - 1.0), rtol=0, atol=0.02)
    assert np.all(np.abs(np.linalg.norm(np.asarray(r["points"]), axis=1) - 1.0) < 0.02)
    assert np.all(np.einsum("ij,ij->i", np.asarray(r["normals"]), d) > 0.99)


def test_closest_points_on_a_square_are_exact(impl):
    sq = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0, 1.0, 0], [0, 1.0, 0]])
    faces = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int64)
    q = np.array([[0.3, 0.6, 1.5], [0.8, 0.2, -0.25], [1.5, 0.25, 0.0]])
    r = impl.closest_points(sq, faces, q)
    np.testing.assert_allclose(np.asarray(r["distances"]), [1.5, 0.25, 0.5], rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.asarray(r["points"]), [[0.3, 0.6, 0], [0.8, 0.2, 0], [1.0, 0.25, 0]], rtol=0,
                               atol=1e-12)
    np.testing.assert_allclose(np.asarray(r["normals"]), np.repeat([[0, 0, 1.0]], 3, 0), rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.asarray(impl.closest_points(sq, faces, sq)["distances"]), 0.0, rtol=0, atol=1e-12)


def test_closest_points_never_lose_to_dense_surface_samples(impl):
    from scipy.spatial import cKDTree

    pts, faces = uv_sphere(rings=8, segments=12)
    u, v = np.meshgrid(np.linspace(0, 1, 17), np.linspace(0, 1, 17))
    keep = (u + v) <= 1.0
    u, v = u[keep], v[keep]
    a, b, c = pts[faces[:, 0]], pts[faces[:, 1]], pts[faces[:, 2]]
    samples = (a[:, None] + (b - a)[:, None] * u[None, :, None] + (c - a)[:, None] * v[None, :, None]).reshape(-1, 3)
    rng = np.random.default_rng(18)
    q = rng.normal(size=(150, 3))
    q *= rng.uniform(0.25, 1.75, size=(150, 1)) / np.linalg.norm(q, axis=1, keepdims=True)
    exact = np.asarray(impl.closest_points(pts, faces, q)["distances"])
    approx, _ = cKDTree(samples).query(q)
    # No sample is closer than the surface; and the samples are no farther
    # than one grid step (the longest edge over 16) from the closest point.
    edges = np.linalg.norm(np.concatenate([b - a, c - b, a - c]), axis=1)
    assert np.all(exact <= approx + 1e-12) and np.all(approx - exact < edges.max() / 16)


def test_closest_points_match_brute_force_over_every_triangle(impl):
    """Exact check: the minimum over all triangles of the FP64 closest point."""
    import nricp_oracle

    pts, faces = uv_sphere(rings=7, segments=9)
    pts = pts + np.random.default_rng(19).normal(scale=0.02, size=pts.shape)
    q = np.random.default_rng(20).normal(size=(60, 3)) * 0.8
    got = impl.closest_points(pts, faces, q)
    best = np.full(len(q), np.inf)
    for t in range(len(faces)):
        one = nricp_oracle.closest_points(pts, faces[t:t + 1], q)
        best = np.minimum(best, np.asarray(one["distances"]))
    assert np.array_equal(np.asarray(got["distances"]), best)


# ------------------------------------------------------ masks and landmarks --

def test_landmark_validation_explains_itself(host_impl):
    t = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    with pytest.raises(ValueError, match="landmark 1.*template vertex 12|template vertex 12"):
        host_impl.validate_landmarks(np.array([2, 12]), t, 10)
    with pytest.raises(ValueError, match="twice"):
        host_impl.validate_landmarks(np.array([4, 4]), t, 10)
    with pytest.raises(ValueError, match="matching indices"):
        host_impl.validate_landmarks(np.array([1, 2]), t[:1], 10)
    host_impl.validate_landmarks(np.array([1, 2]), t, 10)


def test_mask_validation_counts_active_vertices(host_impl):
    m = np.ones(9, dtype=np.int64)
    m[[2, 5]] = 0
    assert host_impl.validate_mask(m, 9) == 7
    with pytest.raises(ValueError, match="mask has 4 entries"):
        host_impl.validate_mask(m[:4], 9)
    with pytest.raises(ValueError, match="excludes every vertex"):
        host_impl.validate_mask(np.zeros(9, dtype=np.int64), 9)
