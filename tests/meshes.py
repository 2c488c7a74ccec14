"""Synthetic fixtures shared by the tests: the reference tests' UV sphere,
smooth displacement, curved patch and seeded graph state
(tests/test_core_geometry.py:26-52, tests/test_route_a.py:18-44,
tests/test_energy_blocks.py:21-130 in the reference), restated."""

from __future__ import annotations

import numpy as np


def uv_sphere(rings: int = 32, segments: int = 64, radius: float = 1.0):
    theta = np.pi * np.arange(rings + 1) / rings
    phi = 2 * np.pi * np.arange(segments) / segments
    t, p = np.meshgrid(theta, phi, indexing="ij")
    vertices = np.stack([radius * np.sin(t) * np.cos(p), radius * np.sin(t) * np.sin(p), radius * np.cos(t)],
                        axis=-1).reshape(-1, 3)
    faces = []
    for ring in range(rings):
        for seg in range(segments):
            a = ring * segments + seg
            b = ring * segments + (seg + 1) % segments
            c = (ring + 1) * segments + (seg + 1) % segments
            d = (ring + 1) * segments + seg
            faces.append([a, c, b])
            faces.append([a, d, c])
    return vertices, np.asarray(faces, dtype=np.int64)


def welded_uv_sphere(rings: int = 32, segments: int = 64, radius: float = 1.0):
    """The UV sphere with its duplicated pole vertices merged and the
    zero-area pole triangles dropped: a mesh on which every equidistant tie
    between triangles shares the tied vertex (and so its normal)."""
    points, faces = uv_sphere(rings, segments, radius)
    remap = np.arange(len(points))
    north, south = 0, rings * segments
    remap[:segments] = north
    remap[rings * segments:] = south
    faces = remap[faces]
    keep = (faces[:, 0] != faces[:, 1]) & (faces[:, 1] != faces[:, 2]) & (faces[:, 0] != faces[:, 2])
    faces = faces[keep]
    used = np.unique(faces)
    compact = -np.ones(len(points), dtype=np.int64)
    compact[used] = np.arange(len(used))
    return points[used], compact[faces]


def bulge(points: np.ndarray, amplitude: float = 0.08) -> np.ndarray:
    radius = np.linalg.norm(points, axis=1, keepdims=True)
    direction = points / np.maximum(radius, 1e-12)
    bump = amplitude * np.exp(-((points[:, 2:3] - 0.3) ** 2) / 0.05)
    out = points + direction * bump
    out[:, 0] += 0.02 * points[:, 2]
    return out


def route_settings(levels: int, **overrides) -> list[dict]:
    base = {
        "point_to_point": 1.0,
        "point_to_plane": 2.0,
        "smoothness": 0.05,
        "anchor": 1e-4,
        "landmark": 10.0,
        "robust_kernel": "huber",
        "robust_threshold": 0.0,
        "max_outer_iterations": 6,
        "max_inner_iterations": 6,
        "outer_tolerance": 1e-4,
        "inner_tolerance": 1e-5,
    }
    base.update(overrides)
    return [dict(base) for _ in range(levels)]


def sphere_patch(scale: float = 1.0, rows: int = 6, columns: int = 7):
    u = np.linspace(-0.5, 0.5, columns)
    v = np.linspace(-0.5, 0.5, rows)
    gu, gv = np.meshgrid(u, v, indexing="ij")
    points = np.stack([gu.ravel() * scale, gv.ravel() * scale, 0.25 * scale * (gu ** 2 - gv ** 2).ravel()], axis=1)
    faces = []
    for i in range(columns - 1):
        for j in range(rows - 1):
            a, b = i * rows + j, (i + 1) * rows + j
            c, d = (i + 1) * rows + j + 1, i * rows + j + 1
            faces += [[a, b, c], [a, c, d]]
    return points, np.asarray(faces, dtype=np.int64)


def rotation_about(axis: np.ndarray, angle: float) -> np.ndarray:
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * k + (1 - np.cos(angle)) * (k @ k)


def graph_state(mod, points: np.ndarray, node_count: int, seed: int) -> dict:
    """A seeded graph, deformation and correspondence set on `points`."""
    rng = np.random.default_rng(seed)
    nodes = points[np.asarray(mod.farthest_point_seeds(points, node_count))]
    binding = mod.bind_vertices(points, nodes)
    rotations = np.stack([rotation_about(rng.normal(size=3), rng.normal(scale=0.05) * 1.7)
                          for _ in range(node_count)])
    translations = rng.normal(scale=0.02, size=(node_count, 3))
    deformed = mod.evaluate_deformation(points, binding["nodes"], binding["weights"], binding["counts"], rotations,
                                        translations, nodes)
    target = deformed + rng.normal(scale=0.01, size=deformed.shape)
    normals = mod.vertex_normals(target, sphere_patch()[1])
    return {
        "nodes": nodes,
        "binding": binding,
        "rotations": rotations,
        "translations": translations,
        "correspondence_points": target,
        "correspondence_normals": normals,
        "edges": np.asarray([[i, (i + 1) % node_count] for i in range(node_count)], dtype=np.int64),
    }


def all_weights() -> dict:
    return {"point_to_point": 1.0, "point_to_plane": 2.0, "smoothness": 0.5, "anchor": 0.25, "landmark": 3.0}


def energy_args(points: np.ndarray, state: dict, accepted=None, robust=None):
    n = len(points)
    return (points, state["binding"]["nodes"], state["binding"]["weights"], state["binding"]["counts"],
            state["rotations"], state["translations"], state["nodes"], state["correspondence_points"],
            state["correspondence_normals"],
            np.ones(n, dtype=np.int64) if accepted is None else accepted,
            np.ones(n) if robust is None else robust, state["edges"],
            np.array([0, n // 2], dtype=np.int64), state["correspondence_points"][[0, n // 2]] + 0.01)
