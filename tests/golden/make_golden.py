"""Writes the golden fixtures under tests/golden/ from the reference's own
CPU code.

The reference's Eigen-free translation units (mesh, sampling, graph, closest
point, correspondence, robust weights, normal-equation assembly, report) are
compiled by oracle/build_oracle.py into oracle/_ref/libnricp_ref.so, behind the
thin C entry points of oracle/ref_shim/ref_capi.cpp. This script feeds them
seeded synthetic inputs and stores inputs and outputs as small .npz / .json
files. The fixtures are data only; the tests compare the oracle (on the CPU)
and the device path (on the GPU) against them.

Run from the repository root after building the reference library:
    python tests/golden/make_golden.py
"""

from __future__ import annotations

import ctypes
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "tests"))

from meshes import bulge, welded_uv_sphere  # noqa: E402

LIB = ROOT / "oracle" / "_ref" / "libnricp_ref.so"

D = ctypes.POINTER(ctypes.c_double)
I32 = ctypes.POINTER(ctypes.c_int32)
I64 = ctypes.POINTER(ctypes.c_int64)
U8 = ctypes.POINTER(ctypes.c_uint8)


def ptr(a, t):
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def load():
    lib = ctypes.CDLL(str(LIB))
    lib.ref_robust_weight.restype = ctypes.c_double
    lib.ref_robust_weight.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double]
    for name in ("ref_clean_mesh", "ref_build_node_edges", "ref_assemble_normal_equations", "ref_sample_report"):
        getattr(lib, name).restype = ctypes.c_int64
    n64, n32, dbl = ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    lib.ref_vertex_normals.argtypes = [D, n64, I32, n64, D]
    lib.ref_clean_mesh.argtypes = [D, n64, I32, n64, I64, I32]
    lib.ref_normalization.argtypes = [D, n64, D, D]
    lib.ref_farthest_point_seeds.argtypes = [D, n64, n32, n32, I32]
    lib.ref_bind_vertices.argtypes = [D, n64, D, n64, n32, I32, D, I32]
    lib.ref_build_node_edges.argtypes = [D, n64, n32, I32]
    lib.ref_closest_points.argtypes = [D, n64, I32, n64, D, n64, D, D, D, I32]
    lib.ref_find_correspondences.argtypes = [D, n64, I32, n64, D, D, n64, dbl, dbl, dbl, U8, I32, I32, D]
    lib.ref_assemble_normal_equations.argtypes = [I32, I32, D, I32, I32, D, n64, n32, D, D, I32, D, n64]
    lib.ref_damp_diagonal.argtypes = [D, n32, dbl, n32]
    lib.ref_sample_report.argtypes = [ctypes.c_char_p, n64, D, I32, D, n32]
    return lib


def jittered_sphere(rings, segments, scale, seed):
    points, faces = welded_uv_sphere(rings=rings, segments=segments)
    rng = np.random.default_rng(seed)
    return points + rng.normal(scale=scale, size=points.shape), faces


def save(name, **arrays):
    np.savez(HERE / f"{name}.npz", **arrays)


def main() -> int:
    lib = load()
    rng = np.random.default_rng(20260)

    # Vertex normals (area-weighted, face order) on a generic closed mesh.
    p, f = jittered_sphere(10, 14, 0.01, 1)
    out = np.zeros_like(p)
    lib.ref_vertex_normals(ptr(f64(p), D), len(p), ptr(i32(f), I32), len(f), ptr(out, D))
    save("vertex_normals", points=p, faces=f, normals=out)

    # Cleaning: a duplicate face, a degenerate face and an isolated vertex.
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [5, 5, 5], [0.5, 0.5, 0.0]], dtype=np.float64)
    f = np.array([[0, 1, 2], [1, 3, 2], [2, 1, 0], [0, 0, 3], [1, 3, 5]], dtype=np.int64)
    pw, fw = f64(p).copy(), i32(f).copy()
    nv, counts = np.zeros(1, dtype=np.int64), np.zeros(3, dtype=np.int32)
    nf = lib.ref_clean_mesh(ptr(pw, D), len(p), ptr(fw, I32), len(f), ptr(nv, I64), ptr(counts, I32))
    save("clean_mesh", points=p, faces=f, clean_points=pw[: nv[0]], clean_faces=fw[:nf].astype(np.int64),
         counts=counts.astype(np.int64))

    # Normalisation of a random cloud.
    p = rng.normal(size=(257, 3)) * np.array([3.0, 1.0, 0.5]) + np.array([10.0, -2.0, 4.0])
    c, s = np.zeros(3), np.zeros(1)
    lib.ref_normalization(ptr(f64(p), D), len(p), ptr(c, D), ptr(s, D))
    save("normalization", points=p, center=c, scale=s)

    # Farthest-point seeds from two starts.
    p = rng.uniform(-1.0, 1.0, size=(300, 3))
    seeds = {}
    for first in (0, 7):
        o = np.zeros(40, dtype=np.int32)
        lib.ref_farthest_point_seeds(ptr(f64(p), D), len(p), 40, first, ptr(o, I32))
        seeds[f"first_{first}"] = o.astype(np.int64)
    save("farthest_point_seeds", points=p, count=np.array(40), **seeds)

    # Bindings: four of thirty nodes, and fewer nodes than the support.
    p = rng.uniform(-1.0, 1.0, size=(200, 3))
    arrays = {"points": p}
    for tag, nodes, k in (("many", p[rng.choice(200, 30, replace=False)], 4), ("few", p[:3], 4),
                          ("three", p[rng.choice(200, 25, replace=False)], 3)):
        on, ow, oc = np.zeros(4 * len(p), np.int32), np.zeros(4 * len(p)), np.zeros(len(p), np.int32)
        lib.ref_bind_vertices(ptr(f64(p), D), len(p), ptr(f64(nodes), D), len(nodes), k, ptr(on, I32), ptr(ow, D),
                              ptr(oc, I32))
        arrays.update({f"{tag}_nodes": nodes, f"{tag}_k": np.array(k), f"{tag}_bound": on.reshape(-1, 4).astype(np.int64),
                       f"{tag}_weights": ow.reshape(-1, 4), f"{tag}_counts": oc.astype(np.int64)})
    save("bind_vertices", **arrays)

    # Node kNN edges for 4 and 8 neighbours.
    nodes = rng.uniform(-1.0, 1.0, size=(50, 3))
    arrays = {"nodes": nodes}
    for k in (4, 8):
        o = np.zeros(2 * len(nodes) * k, np.int32)
        m = lib.ref_build_node_edges(ptr(f64(nodes), D), len(nodes), k, ptr(o, I32))
        arrays[f"edges_{k}"] = o[: 2 * m].reshape(-1, 2).astype(np.int64)
    save("build_node_edges", **arrays)

    # Closest points and correspondences on a bumpy generic target.
    tp, tf = jittered_sphere(12, 16, 0.004, 2)
    tp = bulge(tp, 0.08)
    dirs = rng.normal(size=(400, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    q = dirs * rng.uniform(0.7, 1.3, size=(400, 1))
    op, on_, od, ot = np.zeros_like(q), np.zeros_like(q), np.zeros(len(q)), np.zeros(len(q), np.int32)
    lib.ref_closest_points(ptr(f64(tp), D), len(tp), ptr(i32(tf), I32), len(tf), ptr(f64(q), D), len(q), ptr(op, D),
                           ptr(on_, D), ptr(od, D), ptr(ot, I32))
    save("closest_points", target_points=tp, target_faces=tf, queries=q, points=op, normals=on_, distances=od,
         triangles=ot.astype(np.int64))
    qn = dirs + rng.normal(scale=0.6, size=dirs.shape)
    qn /= np.linalg.norm(qn, axis=1, keepdims=True)
    mask = (rng.uniform(size=len(q)) > 0.1).astype(np.uint8)
    arrays = {"target_points": tp, "target_faces": tf, "queries": q, "query_normals": qn, "mask": mask}
    for tag, (dist, mincos, slope, use_mask) in {"plain": (np.inf, -1.0, 0.0, False),
                                                  "filtered": (0.15, 0.3, 0.2, True)}.items():
        rej, cnt, stats = np.zeros(len(q), np.int32), np.zeros(6, np.int32), np.zeros(3)
        lib.ref_find_correspondences(ptr(f64(tp), D), len(tp), ptr(i32(tf), I32), len(tf), ptr(f64(q), D),
                                     ptr(f64(qn), D), len(q), dist, mincos, slope,
                                     ptr(mask, U8) if use_mask else None, ptr(rej, I32), ptr(cnt, I32), ptr(stats, D))
        arrays.update({f"{tag}_settings": np.array([dist, mincos, slope, float(use_mask)]),
                       f"{tag}_rejection": rej.astype(np.int64), f"{tag}_counts": cnt.astype(np.int64),
                       f"{tag}_stats": stats})
    save("find_correspondences", **arrays)

    # Robust weights of both kernels over a range of residuals and thresholds.
    sq = np.concatenate([[0.0], np.geomspace(1e-8, 10.0, 40)])
    arrays = {"squared": sq}
    for kernel in ("huber", "geman_mcclure"):
        for t in (0.0, 0.05, 0.5):
            arrays[f"{kernel}_{t}"] = np.array([lib.ref_robust_weight(kernel.encode(), float(s), t) for s in sq])
    save("robust_weights", **arrays)

    # Normal-equation assembly from random residual blocks.
    nn, nb = 9, 60
    rows_of = {0: 3, 1: 1, 2: 3, 3: 6, 4: 3}
    kind = rng.integers(0, 5, size=nb)
    rows = np.array([rows_of[int(k)] for k in kind])
    slots = np.array([1 if k == 3 else (2 if k == 2 else int(rng.integers(1, 4))) for k in kind])
    nodes = np.full((nb, 4), -1, dtype=np.int64)
    for i in range(nb):
        nodes[i, : slots[i]] = rng.choice(nn, slots[i], replace=False)
    values = np.zeros((nb, 6))
    jac = np.zeros((nb, 4, 6, 6))
    for i in range(nb):
        values[i, : rows[i]] = rng.normal(size=rows[i])
        jac[i, : slots[i], : rows[i], :] = rng.normal(size=(slots[i], rows[i], 6))
    diag, rhs = np.zeros((nn, 36)), np.zeros((nn, 6))
    cap = nn * nn
    pairs, blocks = np.zeros(2 * cap, np.int32), np.zeros(36 * cap)
    npairs = lib.ref_assemble_normal_equations(ptr(i32(kind), I32), ptr(i32(rows), I32), ptr(f64(values), D),
                                               ptr(i32(slots), I32), ptr(i32(nodes), I32), ptr(f64(jac), D), nb, nn,
                                               ptr(diag, D), ptr(rhs, D), ptr(pairs, I32), ptr(blocks, D), cap)
    damped = {}
    for lam, scaling in ((0.0, 0), (0.3, 1), (0.3, 0)):
        d = diag.copy()
        lib.ref_damp_diagonal(ptr(d, D), nn, lam, scaling)
        damped[f"damped_{lam}_{scaling}"] = d.reshape(nn, 6, 6)
    save("assemble_normal_equations", kind=kind, rows=rows, values=values, slots=slots, nodes=nodes, jacobians=jac,
         node_count=np.array(nn), diagonal=diag.reshape(nn, 6, 6), rhs=rhs,
         pairs=pairs[: 2 * npairs].reshape(-1, 2).astype(np.int64), blocks=blocks[: 36 * npairs].reshape(-1, 6, 6),
         **damped)

    # A report with every field set, and one with non-finite numbers.
    cases = []
    for energies, counts, stats, levels in (
            ([0.5, 0.25, 0.125, 1e-3, 0.0], [100, 90, 4, 3, 2, 1], [0.01, 0.05, 0.2], 3),
            ([np.inf, np.nan, 1.0, 2.0, 3.0], [10, 0, 10, 0, 0, 0], [0.0, 0.0, 0.0], 0)):
        buf = ctypes.create_string_buffer(1 << 16)
        n = lib.ref_sample_report(buf, 1 << 16, ptr(f64(energies), D), ptr(i32(counts), I32), ptr(f64(stats), D),
                                  levels)
        cases.append({"energies": [float(e) if np.isfinite(e) else str(e) for e in energies], "counts": counts,
                      "stats": stats, "levels": levels, "json": buf.value[:n].decode()})
    (HERE / "report.json").write_text(json.dumps(cases, indent=1) + "\n")
    print(f"fixtures written to {HERE}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
