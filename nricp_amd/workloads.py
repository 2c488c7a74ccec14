"""Seeded synthetic wrap jobs at the benchmark shapes.

No dataset is available, so every workload is generated: a class-I geodesic
sphere (10 f^2 + 2 vertices, 20 f^2 triangles), Gaussian RBF bumps along the
normals, vertex jitter, triangle shuffles, and a capsule-union "humanoid"
meshed by marching tetrahedra. All randomness goes through one
``numpy.random.Generator`` per job so a seed names a job exactly.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_PHI = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_VERTS = np.array(
    [[-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
     [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
     [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1]], dtype=np.float64)
_ICO_FACES = np.array(
    [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
     [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
     [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
     [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)


def geodesic_sphere(frequency: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit geodesic sphere of class I with the given frequency.

    Every icosahedron face is split into f^2 triangles on its barycentric grid
    and the grid points are projected to the sphere; points shared by faces are
    merged through exact integer keys (corner index + grid coordinates), so the
    result has exactly 10 f^2 + 2 vertices and 20 f^2 outward-wound triangles.
    """
    f = int(frequency)
    if f < 1:
        raise ValueError("frequency must be >= 1")
    corners = _ICO_VERTS / np.linalg.norm(_ICO_VERTS, axis=1, keepdims=True)
    keys: dict[tuple, int] = {}
    points: list[np.ndarray] = []

    def key_of(face: np.ndarray, i: int, j: int) -> tuple:
        # Barycentric integer weights on the face corners; a point is keyed by
        # its (corner, weight) pairs with zero weights dropped, sorted, so the
        # same point on a shared edge or vertex gets the same key.
        k = f - i - j
        parts = sorted((int(c), w) for c, w in ((face[0], k), (face[1], i), (face[2], j)) if w > 0)
        return tuple(parts)

    faces_out = []
    for face in _ICO_FACES:
        a, b, c = corners[face[0]], corners[face[1]], corners[face[2]]
        index = {}
        for i in range(f + 1):
            for j in range(f + 1 - i):
                key = key_of(face, i, j)
                if key not in keys:
                    k = f - i - j
                    p = (a * k + b * i + c * j) / f
                    keys[key] = len(points)
                    points.append(p / np.linalg.norm(p))
                index[(i, j)] = keys[key]
        for i in range(f):
            for j in range(f - i):
                faces_out.append((index[(i, j)], index[(i + 1, j)], index[(i, j + 1)]))
                if i + j + 1 < f:
                    faces_out.append((index[(i + 1, j)], index[(i + 1, j + 1)], index[(i, j + 1)]))
    return np.asarray(points, dtype=np.float64), np.asarray(faces_out, dtype=np.int64)


def frequency_for_vertices(vertex_count: int) -> int:
    """Frequency whose geodesic sphere has about ``vertex_count`` vertices."""
    return max(1, int(round(math.sqrt(max(vertex_count - 2, 10) / 10.0))))


def bbox_diagonal(points: np.ndarray) -> float:
    return float(np.linalg.norm(points.max(axis=0) - points.min(axis=0)))


def vertex_normals(points: np.ndarray, faces: np.ndarray) -> np.ndarray:
    a, b, c = points[faces[:, 0]], points[faces[:, 1]], points[faces[:, 2]]
    w = np.cross(b - a, c - a)
    n = np.zeros_like(points)
    for k in range(3):
        np.add.at(n, faces[:, k], w)
    length = np.linalg.norm(n, axis=1, keepdims=True)
    return n / np.maximum(length, 1e-300)


@dataclass
class BumpField:
    """``count`` Gaussian bumps on the sphere of directions around ``origin``:
    displacement(x) = sum_k amp_k exp(-(angle(x, centre_k) / width_k)^2)."""
    origin: np.ndarray
    centres: np.ndarray
    amps: np.ndarray
    widths: np.ndarray

    def displacement(self, points: np.ndarray) -> np.ndarray:
        dirs = points - self.origin
        dirs = dirs / np.maximum(np.linalg.norm(dirs, axis=1, keepdims=True), 1e-300)
        angle = np.arccos(np.clip(dirs @ self.centres.T, -1.0, 1.0))
        return (self.amps[None, :] * np.exp(-(angle / self.widths[None, :]) ** 2)).sum(axis=1)


def bump_field(points: np.ndarray, rng: np.random.Generator, count: int = 8,
               amplitude: tuple[float, float] = (0.05, 0.15), width: tuple[float, float] = (0.2, 0.5),
               center: np.ndarray | None = None) -> BumpField:
    """Amplitude U(a) x the points' bbox diagonal, angular width U(w) radians,
    centres uniform on the sphere."""
    diag = bbox_diagonal(points)
    origin = points.mean(axis=0) if center is None else center
    centres = rng.normal(size=(count, 3))
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    amps = rng.uniform(*amplitude, size=count) * diag
    widths = rng.uniform(*width, size=count)
    return BumpField(np.asarray(origin, dtype=np.float64), centres, amps, widths)


def rbf_bumps(points: np.ndarray, normals: np.ndarray, rng: np.random.Generator, count: int = 8,
              amplitude: tuple[float, float] = (0.05, 0.15), width: tuple[float, float] = (0.2, 0.5),
              center: np.ndarray | None = None) -> np.ndarray:
    """Displace along the normals by a bump field (bump_field)."""
    field = bump_field(points, rng, count, amplitude, width, center)
    return points + normals * field.displacement(points)[:, None]


def shuffle_triangles(faces: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    return faces[rng.permutation(len(faces))]


# ------------------------------------------------------------- capsule body --

@dataclass
class CapsuleBody:
    """A chain-linked set of capsules: segment endpoints, radii and the joint
    (parent endpoint) each segment hangs from."""
    starts: np.ndarray
    ends: np.ndarray
    radii: np.ndarray
    parents: list[int] = field(default_factory=list)
    motion: tuple | None = None  # set by posed(): (rotations, rest starts, posed starts)

    def posed(self, rng: np.random.Generator, max_degrees: float = 30.0) -> "CapsuleBody":
        """Each segment rotated rigidly about its own joint (its start point) by
        up to ``max_degrees``; children follow their parent's end point."""
        starts = self.starts.copy()
        ends = self.ends.copy()
        rotations = []
        for i in range(len(starts)):
            p = self.parents[i]
            if p >= 0:
                shift = ends[p] - self.ends[p] if self.starts[i] is not None else 0.0
                starts[i] = self.starts[i] + shift
                ends[i] = self.ends[i] + shift
            axis = rng.normal(size=3)
            axis /= np.linalg.norm(axis)
            angle = np.deg2rad(rng.uniform(-max_degrees, max_degrees))
            k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
            r = np.eye(3) + np.sin(angle) * k + (1 - np.cos(angle)) * (k @ k)
            ends[i] = starts[i] + r @ (ends[i] - starts[i])
            rotations.append(r)
        posed = CapsuleBody(starts, ends, self.radii.copy(), list(self.parents))
        # Segment i maps x to R_i (x - start_i) + start'_i.
        posed.motion = (np.array(rotations), self.starts.copy(), starts.copy())
        return posed

    def nearest_segment(self, x: np.ndarray) -> np.ndarray:
        """Index of the capsule whose surface is nearest to each point."""
        best = None
        index = np.zeros(len(x), dtype=np.int64)
        for i, (s, e, r) in enumerate(zip(self.starts, self.ends, self.radii)):
            d = e - s
            t = np.clip(((x - s) @ d) / max(float(d @ d), 1e-12), 0.0, 1.0)
            dist = np.abs(np.linalg.norm(x - (s + t[:, None] * d), axis=1) - r)
            if best is None:
                best = dist
            else:
                closer = dist < best
                best = np.where(closer, dist, best)
                index = np.where(closer, i, index)
        return index

    def sdf(self, x: np.ndarray) -> np.ndarray:
        """Smooth-union signed distance (polynomial smooth min, k = 0.02)."""
        out = None
        for s, e, r in zip(self.starts, self.ends, self.radii):
            d = e - s
            t = np.clip(((x - s) @ d) / max(float(d @ d), 1e-12), 0.0, 1.0)
            dist = np.linalg.norm(x - (s + t[:, None] * d), axis=1) - r
            if out is None:
                out = dist
            else:
                k = 0.02
                h = np.clip(0.5 + 0.5 * (dist - out) / k, 0.0, 1.0)
                out = dist * (1 - h) + out * h - k * h * (1 - h)
        return out


def humanoid(rng: np.random.Generator, segments: int = 15) -> CapsuleBody:
    """15 capsules (radius U(0.03, 0.12), length U(0.1, 0.4)) chained at joints:
    a torso chain with limbs branching from it."""
    starts, ends, radii, parents = [], [], [], []
    directions = [np.array([0, 0, 1.0]), np.array([0, 0, -1.0]), np.array([1.0, 0, 0]),
                  np.array([-1.0, 0, 0])]
    for i in range(segments):
        radius = rng.uniform(0.03, 0.12)
        length = rng.uniform(0.1, 0.4)
        if i == 0:
            parent, start, direction = -1, np.zeros(3), np.array([0, 0, 1.0])
        else:
            parent = int(rng.integers(0, i)) if i > 2 else i - 1
            start = ends[parent]
            direction = directions[i % 4] + rng.normal(scale=0.35, size=3)
            direction /= np.linalg.norm(direction)
        starts.append(start)
        ends.append(start + direction * length)
        radii.append(radius)
        parents.append(parent)
    return CapsuleBody(np.array(starts), np.array(ends), np.array(radii), parents)


# The six tetrahedra of a cube along its main diagonal (corner 0 -> corner 6),
# corners numbered as (x, y, z) bits: 0=(000) 1=(100) 2=(110) 3=(010)
# 4=(001) 5=(101) 6=(111) 7=(011).
_CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
_TETS = np.array([[0, 5, 1, 6], [0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6], [0, 4, 5, 6]])


def marching_tetrahedra(sdf, lo: np.ndarray, hi: np.ndarray, resolution: int) -> tuple[np.ndarray, np.ndarray]:
    """Zero level set of ``sdf`` on a regular grid, split into tetrahedra.

    Vertices sit on grid edges (linear interpolation) and are shared through
    the grid-edge key, so the output is a closed manifold wherever the level
    set is; triangles are wound with the outward (positive) side.
    """
    n = int(resolution)
    axes = [np.linspace(lo[k], hi[k], n + 1) for k in range(3)]
    gx, gy, gz = np.meshgrid(*axes, indexing="ij")
    grid = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    values = sdf(grid).reshape(n + 1, n + 1, n + 1)
    flat = values.ravel()
    cell = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), axis=-1).reshape(-1, 3)
    corner_ids = []
    for c in _CUBE:
        p = cell + c
        corner_ids.append((p[:, 0] * (n + 1) + p[:, 1]) * (n + 1) + p[:, 2])
    corner_ids = np.stack(corner_ids, axis=1)
    keep = (np.sign(flat[corner_ids]) != np.sign(flat[corner_ids[:, :1]])).any(axis=1)
    corner_ids = corner_ids[keep]
    tris = []
    edge_vertex: dict[tuple[int, int], int] = {}
    verts: list[np.ndarray] = []

    def vertex_on(a: int, b: int) -> int:
        key = (a, b) if a < b else (b, a)
        idx = edge_vertex.get(key)
        if idx is None:
            va, vb = flat[a], flat[b]
            t = va / (va - vb)
            idx = len(verts)
            verts.append(grid[a] + t * (grid[b] - grid[a]))
            edge_vertex[key] = idx
        return idx

    for tet in _TETS:
        ids = corner_ids[:, tet]
        inside = flat[ids] < 0.0
        count = inside.sum(axis=1)
        for row, cnt in zip(np.nonzero((count > 0) & (count < 4))[0], count[(count > 0) & (count < 4)]):
            v = ids[row]
            ins = inside[row]
            inner = [v[k] for k in range(4) if ins[k]]
            outer = [v[k] for k in range(4) if not ins[k]]
            if cnt == 1 or cnt == 3:
                lone = inner[0] if cnt == 1 else outer[0]
                others = outer if cnt == 1 else inner
                tri = [vertex_on(lone, o) for o in others]
                tris.append(tri)
            else:
                a, b = inner
                c, d = outer
                q = [vertex_on(a, c), vertex_on(a, d), vertex_on(b, d), vertex_on(b, c)]
                tris.append([q[0], q[1], q[2]])
                tris.append([q[0], q[2], q[3]])
    points = np.asarray(verts)
    faces = np.asarray(tris, dtype=np.int64)
    # Orient every triangle so its normal points along the SDF gradient.
    a, b, c = points[faces[:, 0]], points[faces[:, 1]], points[faces[:, 2]]
    normal = np.cross(b - a, c - a)
    centre = (a + b + c) / 3.0
    eps = 1e-4 * float(np.linalg.norm(hi - lo))
    grad = np.stack([(sdf(centre + eps * e) - sdf(centre - eps * e)) for e in np.eye(3)], axis=1)
    flip = np.einsum("ij,ij->i", normal, grad) < 0.0
    faces[flip] = faces[flip][:, ::-1]
    # Drop degenerate slivers and unreferenced vertices.
    a, b, c = points[faces[:, 0]], points[faces[:, 1]], points[faces[:, 2]]
    area = np.linalg.norm(np.cross(b - a, c - a), axis=1)
    faces = faces[area > 1e-12 * float(np.linalg.norm(hi - lo)) ** 2]
    used = np.unique(faces)
    remap = -np.ones(len(points), dtype=np.int64)
    remap[used] = np.arange(len(used))
    return points[used], remap[faces]


def mesh_body(body: CapsuleBody, target_vertices: int) -> tuple[np.ndarray, np.ndarray]:
    """Mesh a capsule body to roughly ``target_vertices`` vertices by picking
    the grid resolution from the surface-area scaling (two trial meshes)."""
    lo = np.minimum(body.starts, body.ends).min(axis=0) - body.radii.max() - 0.05
    hi = np.maximum(body.starts, body.ends).max(axis=0) + body.radii.max() + 0.05
    res = 40
    points, faces = marching_tetrahedra(body.sdf, lo, hi, res)
    res = max(8, int(round(res * math.sqrt(target_vertices / max(len(points), 1)))))
    return marching_tetrahedra(body.sdf, lo, hi, res)


# ------------------------------------------------------------------ jobs ----

@dataclass
class WrapJob:
    name: str
    template_points: np.ndarray
    template_faces: np.ndarray
    target_points: np.ndarray
    target_faces: np.ndarray
    nodes_final: int
    graph_levels: int
    seed: int
    # Where each template vertex belongs on the target, when the generator
    # knows it (the quality harness's correspondence error); None otherwise.
    ground_truth: np.ndarray | None = None

    @property
    def nodes_initial(self) -> int:
        return max(1, int(round(self.nodes_final / 2 ** (self.graph_levels - 1))))


def sphere_job(template_frequency: int, target_frequency: int, seed: int, nodes_final: int,
               graph_levels: int, jitter: float = 1e-3, normal_noise: float = 0.0,
               name: str = "sphere") -> WrapJob:
    """Geodesic-sphere template; a denser geodesic sphere with 8 RBF bumps,
    vertex jitter (fraction of the diagonal) or normal noise, and shuffled
    triangles as the target."""
    rng = np.random.default_rng(seed)
    tp, tf = geodesic_sphere(template_frequency)
    gp, gf = geodesic_sphere(target_frequency)
    gn = vertex_normals(gp, gf)
    field = bump_field(gp, rng)
    gp = gp + gn * field.displacement(gp)[:, None]
    # The clean target surface in each template vertex's direction (the unit
    # sphere displaced radially by the bump field).
    truth = tp / np.linalg.norm(tp, axis=1, keepdims=True)
    truth = truth * (1.0 + field.displacement(tp))[:, None]
    diag = bbox_diagonal(gp)
    if jitter > 0.0:
        gp = gp + rng.normal(scale=jitter * diag, size=gp.shape)
    if normal_noise > 0.0:
        gp = gp + vertex_normals(gp, gf) * rng.normal(scale=normal_noise * diag, size=(len(gp), 1))
    gf = shuffle_triangles(gf, rng)
    return WrapJob(name, tp, tf, gp, gf, nodes_final, graph_levels, seed, truth)


def humanoid_job(seed: int, template_vertices: int = 50_000, target_triangles: int = 120_000,
                 nodes_final: int = 2000, graph_levels: int = 4, hole_fraction: float = 0.02,
                 normal_noise: float = 0.0, name: str = "humanoid") -> WrapJob:
    """Capsule humanoid template; posed + bumped + independently meshed target
    with a fraction of its triangles deleted, and optionally Gaussian noise
    along the target's normals (sigma a fraction of its diagonal)."""
    rng = np.random.default_rng(seed)
    body = humanoid(rng)
    tp, tf = mesh_body(body, template_vertices)
    posed = body.posed(rng)
    # Each template vertex carried by its nearest capsule's rigid motion (the
    # target's small bumps and holes aside).
    rotations, rest_starts, posed_starts = posed.motion
    seg = body.nearest_segment(tp)
    truth = np.einsum("nij,nj->ni", rotations[seg], tp - rest_starts[seg]) + posed_starts[seg]
    gp, gf = mesh_body(posed, target_triangles // 2)
    gp = rbf_bumps(gp, vertex_normals(gp, gf), rng, amplitude=(0.005, 0.015))
    if normal_noise > 0.0:
        diag = bbox_diagonal(gp)
        gp = gp + vertex_normals(gp, gf) * rng.normal(scale=normal_noise * diag, size=(len(gp), 1))
    keep = rng.random(len(gf)) >= hole_fraction
    gf = shuffle_triangles(gf[keep], rng)
    used = np.unique(gf)
    remap = -np.ones(len(gp), dtype=np.int64)
    remap[used] = np.arange(len(used))
    return WrapJob(name, tp, tf, gp[used], remap[gf], nodes_final, graph_levels, seed, truth)


def level_settings(levels: int, outer: int = 10, inner: int = 1, **weights) -> list[dict]:
    """The benchmark schedule: ``outer`` correspondence updates of ``inner``
    Gauss-Newton steps per level, tolerances zero so every run does the same
    number of iterations."""
    base = {
        "point_to_point": 0.1,
        "point_to_plane": 1.0,
        "smoothness": 1.0,
        "anchor": 1e-4,
        "landmark": 10.0,
        "robust_kernel": "huber",
        "robust_threshold": 0.0,
        "max_outer_iterations": outer,
        "max_inner_iterations": inner,
        "outer_tolerance": 0.0,
        "inner_tolerance": 0.0,
    }
    base.update(weights)
    return [dict(base) for _ in range(levels)]


def config_job(index: int, seed: int = 0) -> WrapJob:
    """The benchmark configurations by index (0: CPU case, 1: humanoid 50k,
    2: 200k single job on geodesic spheres, 3: the 200k single job with the
    humanoid generator of configuration 1)."""
    if index == 0:
        return sphere_job(32, 64, seed, nodes_final=500, graph_levels=3, name="icosphere-10k")
    if index == 1:
        return humanoid_job(seed)
    if index == 2:
        return sphere_job(141, 200, seed, nodes_final=8000, graph_levels=3, jitter=0.0,
                          normal_noise=5e-4, name="geodesic-200k")
    if index == 3:
        return humanoid_job(seed, template_vertices=200_000, target_triangles=800_000, nodes_final=8000,
                            graph_levels=3, normal_noise=5e-4, name="humanoid-200k")
    raise ValueError(f"no single-job configuration {index}")


def batch_jobs(count: int = 64, seed: int = 1000, lo: int = 2000, hi: int = 50_000,
               kind: str = "humanoid", indices=None) -> list[WrapJob]:
    """Uneven batch: template sizes log-uniform in [lo, hi], target 2-3x the
    template's triangles, nodes ~ vertices / 25, one seed per job. "humanoid"
    generates each job as the 50k humanoid configuration does (4 levels);
    "sphere" uses geodesic spheres (3 levels), which are much faster to
    generate. ``indices`` selects jobs of the batch by position: every job's
    size and seed are drawn first, so a selected job is the same job as in the
    whole batch."""
    rng = np.random.default_rng(seed)
    sizes = []
    for _ in range(count):
        v = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        sizes.append((v, rng.uniform(2.0, 3.0)))
    jobs = []
    for k in (range(count) if indices is None else indices):
        v, ratio = sizes[k]
        if kind == "humanoid":
            nodes = max(8, int(round(v / 25)))
            job = humanoid_job(seed + k, template_vertices=int(v), target_triangles=int(2 * v * ratio),
                               nodes_final=nodes, graph_levels=4)
            job.name = f"batch{k}"
            jobs.append(job)
        elif kind == "sphere":
            f = frequency_for_vertices(int(v))
            g = max(f + 1, int(round(f * math.sqrt(ratio))))
            nodes = max(8, int(round((10 * f * f + 2) / 25)))
            jobs.append(sphere_job(f, g, seed + k, nodes_final=nodes, graph_levels=3, name=f"batch{k}"))
        else:
            raise ValueError(f"unknown batch kind {kind!r}")
    return jobs


def batch_sizes(count: int = 64, seed: int = 1000, lo: int = 2000, hi: int = 50_000) -> list[float]:
    """The template vertex counts batch_jobs draws, in job order."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        out.append(float(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        rng.uniform(2.0, 3.0)
    return out
