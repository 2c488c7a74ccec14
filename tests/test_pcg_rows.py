"""The device PCG on block rows of every length its matrix product handles.

Each node's row is multiplied by its eight lanes, lane j taking blocks j,
j + 8, j + 16, ...: rows of at most 8 blocks use one block per lane, rows of
9 to 16 a second one (whose column an FP32 matrix keeps in a register), and
longer rows read the remaining columns from memory. The systems below hold
all three kinds, at sizes that select the single-group kernel (one and two
node slots per lane group) and the grid kernels (one and two slots). Every
solve is checked against a sparse direct solve of the damped system it
solves: the FP64 and mixed forms the given system (the mixed form iterates on
the matrix rounded to FP32 and refines against the FP64 one), the all-FP32
form the system whose matrix and damping are rounded to FP32 (its FP64
refinement drives the true residual of that system to the tolerance).
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu

DAMPING = 1e-3


def random_system(n: int, seed: int):
    """A block-diagonally dominant (so SPD) system: a ring, three edges per
    node to random nodes at most eight positions on, and hubs joined to 24
    random nodes at most 64 positions on. Every edge is local, so the direct
    solve's factor stays sparse."""
    rng = np.random.default_rng(seed)
    edges = {(min(i, (i + 1) % n), max(i, (i + 1) % n)) for i in range(n)}
    for i in range(n):
        for d in rng.choice(np.arange(2, 9), size=3, replace=False):
            j = (i + int(d)) % n
            if j != i:
                edges.add((min(i, j), max(i, j)))
    for h in rng.choice(n, size=max(1, n // 50), replace=False):
        for d in rng.choice(np.arange(1, 65), size=24, replace=False):
            j = (h + int(d)) % n
            if j != h:
                edges.add((min(h, j), max(h, j)))
    edges = np.array(sorted(edges), dtype=np.int32)
    blocks = rng.normal(scale=0.1, size=(len(edges), 6, 6))
    degree = np.bincount(edges.reshape(-1), minlength=n)
    m = rng.normal(scale=0.1, size=(n, 6, 6))
    diag = (m + m.transpose(0, 2, 1)) / 2 + (1.0 + degree)[:, None, None] * np.eye(6)
    rhs = rng.normal(size=(n, 6))
    return diag, edges, blocks, rhs, degree


def system_matrix(diag, edges, blocks, damping):
    n = len(diag)
    rows, cols, vals = [], [], []
    r6, c6 = np.meshgrid(np.arange(6), np.arange(6), indexing="ij")

    def put(a, b, blk):
        rows.append((6 * a + r6).ravel())
        cols.append((6 * b + c6).ravel())
        vals.append(blk.ravel())

    for i in range(n):
        put(i, i, diag[i] + damping * np.eye(6))
    for (a, b), blk in zip(edges, blocks):
        put(a, b, blk)
        put(b, a, blk.T)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(6 * n, 6 * n))


def direct_solve(diag, edges, blocks, rhs, damping):
    return spla.spsolve(system_matrix(diag, edges, blocks, damping).tocsc(), rhs.reshape(-1))


def block_jacobi_pcg_iterations(diag, edges, blocks, rhs, damping, tol):
    """Iterations of the reference's PCG (block-Jacobi, x0 = 0, stop when the
    recurrence residual reaches tol |b|) in numpy, FP64."""
    A = system_matrix(diag, edges, blocks, damping)
    minv = np.linalg.inv(diag + damping * np.eye(6))
    precondition = lambda r: np.einsum("nij,nj->ni", minv, r.reshape(-1, 6)).reshape(-1)  # noqa: E731
    b = rhs.reshape(-1)
    r = b.copy()
    z = precondition(r)
    p = z.copy()
    rz = r @ z
    for it in range(10_000):
        if np.linalg.norm(r) <= tol * np.linalg.norm(b):
            return it
        q = A @ p
        alpha = rz / (p @ q)
        r = r - alpha * q
        z = precondition(r)
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    return 10_000


# 40 and 100 nodes run in one workgroup (one and two node slots per lane
# group); 1,500 and 3,000 take the grid kernels (24 groups of one slot, 24
# groups of two); 8,000 is the 200k-vertex job's finest level (63 groups of
# two slots); 20,000 is above what one persistent launch holds (16,384) and
# runs as the multi-launch PCG.
@pytest.mark.parametrize("n", [40, 100, 1500, 3000, 8000, 20_000])
@pytest.mark.parametrize("precision", ["double", "mixed", "float"])
def test_rows_of_every_length_match_a_direct_solve(device, n, precision):
    diag, edges, blocks, rhs, degree = random_system(n, seed=n)
    row_blocks = degree + 1
    assert (row_blocks <= 8).any() and ((row_blocks > 8) & (row_blocks <= 16)).any() and (row_blocks > 16).any()
    if precision != "float":
        expected = direct_solve(diag, edges, blocks, rhs, DAMPING)
    else:
        f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)  # noqa: E731
        expected = direct_solve(f32(diag), edges, f32(blocks), rhs, float(np.float32(DAMPING)))
    r = device.solve_normal_equations(diag, edges, blocks, rhs, damping=DAMPING, precision=precision)
    assert r["ok"]
    got = np.asarray(r["delta"]).reshape(-1)
    np.testing.assert_allclose(got, expected, rtol=0, atol=1e-10 * np.abs(expected).max())
    # The reported residual is the given FP64 system's: for the all-FP32 form
    # it is the FP32 rounding of the matrix (about 6e-8 relative) times |x|.
    bound = 1e-6 if precision == "float" else 1e-11
    assert r["residual_norm"] <= bound * np.linalg.norm(rhs)
    if n > 16_384 and precision != "float":
        # The multi-launch PCG is the reference's recurrence in FP64 (the
        # mixed form solves its FP64 system there): the same iteration count
        # as a numpy block-Jacobi PCG, up to rounding at the stopping test.
        expected_it = block_jacobi_pcg_iterations(diag, edges, blocks, rhs, DAMPING, 1e-13)
        assert abs(r["iterations"] - expected_it) <= 1, (r["iterations"], expected_it)
