"""The allocation-free solve (nricp_hip_solve_bsr_noalloc) against the
allocating one (nricp_hip_solve_system), both through the C ABI.

The allocation-free entry point takes the system already on the device in
block-sparse row form and keeps its workspace in the library's static
per-device array. Given the same blocks in the row order the allocating entry
point builds (each row's blocks by ascending column), the two run the same
kernels on the same values, so their solutions and outcomes agree bit for
bit. Two calls in a row, the second smaller, check that the array's contents
from the first call do not leak into the second; sizes beyond the array's
stated limits are rejected with an error.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
LIBRARY = ROOT / "nricp_amd" / "libnricp_hip.so"
PRECISION = {"double": 0, "mixed": 1, "float": 2}
DAMPING = 1e-3


def random_system(n: int, seed: int):
    """A block-diagonally dominant (so SPD) system: a ring plus three local
    edges per node (rows of up to about eight blocks, some longer)."""
    rng = np.random.default_rng(seed)
    edges = {(min(i, (i + 1) % n), max(i, (i + 1) % n)) for i in range(n)}
    for i in range(n):
        for d in rng.choice(np.arange(2, 12), size=3, replace=False):
            j = (i + int(d)) % n
            if j != i:
                edges.add((min(i, j), max(i, j)))
    edges = np.array(sorted(edges), dtype=np.int32)
    blocks = rng.normal(scale=0.1, size=(len(edges), 6, 6))
    degree = np.bincount(edges.reshape(-1), minlength=n)
    m = rng.normal(scale=0.1, size=(n, 6, 6))
    diag = (m + m.transpose(0, 2, 1)) / 2 + (1.0 + degree)[:, None, None] * np.eye(6)
    rhs = rng.normal(size=(n, 6))
    return diag, edges, blocks, rhs


def bsr_of(diag, edges, blocks):
    """Rows of the full symmetric matrix, each row's blocks by ascending
    column (the order nricp_hip_solve_system builds)."""
    n = len(diag)
    rows = [[(i, diag[i])] for i in range(n)]
    for (a, b), blk in zip(edges, blocks):
        rows[a].append((int(b), blk))
        rows[b].append((int(a), blk.T))
    offsets, columns, values, diag_pos = [0], [], [], np.zeros(n, dtype=np.int32)
    for i, row in enumerate(rows):
        row.sort(key=lambda e: e[0])
        for col, blk in row:
            if col == i:
                diag_pos[i] = len(columns)
            columns.append(col)
            values.append(np.ascontiguousarray(blk))
        offsets.append(len(columns))
    return (np.asarray(offsets, dtype=np.int32), np.asarray(columns, dtype=np.int32), diag_pos,
            np.ascontiguousarray(np.stack(values)))


@pytest.fixture(scope="module")
def lib():
    library = ctypes.CDLL(str(LIBRARY))
    library.nricp_hip_last_error.restype = ctypes.c_char_p
    p, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    library.nricp_hip_solve_system.argtypes = [p, i64, p, p, i64, p, dbl, i32, i32, dbl, i32, p, p]
    library.nricp_hip_solve_bsr_noalloc.argtypes = [p, p, p, i64, i64, p, p, dbl, i32, i32, dbl, i32, p, p, p]
    return library


class Hip:
    """The caller's side: device arrays and a stream from the HIP runtime the
    library uses (the test allocates; the entry point under test does not)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.owned = []

    def check(self, status, what):
        assert status == 0, f"{what} failed with HIP status {status}"

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self.check(self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 1))), "hipMalloc")
        self.owned.append(p)
        self.check(self.rt.hipMemcpy(p, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1), "hipMemcpy")
        return p.value

    def download(self, address: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        self.check(self.rt.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(address),
                                     ctypes.c_size_t(out.nbytes), 2), "hipMemcpy")
        return out

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self.check(self.rt.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def close(self, stream=None):
        self.check(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        for p in self.owned:
            self.rt.hipFree(p)
        self.owned.clear()
        if stream:
            self.rt.hipStreamDestroy(ctypes.c_void_p(stream))


@pytest.fixture()
def hip():
    h = Hip()
    yield h
    h.close()


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def allocating_solve(lib, diag, edges, blocks, rhs, precision):
    n = len(diag)
    delta, info = np.zeros((n, 6)), np.zeros(4)
    d, e, b, r = (np.ascontiguousarray(x) for x in (diag, edges.astype(np.int32), blocks, rhs))
    status = lib.nricp_hip_solve_system(ptr(d), n, ptr(e), ptr(b), len(edges), ptr(r), DAMPING, 1,
                                        PRECISION[precision], 1e-12, 5000, ptr(delta), ptr(info))
    assert status == 0, lib.nricp_hip_last_error()
    return delta, info


def noalloc_solve(lib, hip, diag, edges, blocks, rhs, precision, stream):
    offsets, columns, diag_pos, values = bsr_of(diag, edges, blocks)
    d_off, d_col, d_pos, d_val, d_rhs = (hip.upload(x) for x in (offsets, columns, diag_pos, values, rhs))
    like = np.zeros((len(diag), 6))
    d_delta = hip.upload(like)
    info = np.zeros(4)
    status = lib.nricp_hip_solve_bsr_noalloc(d_off, d_col, d_pos, len(diag), len(columns), d_val, d_rhs, DAMPING, 1,
                                             PRECISION[precision], 1e-12, 5000, d_delta, ptr(info), stream)
    assert status == 0, lib.nricp_hip_last_error()
    return hip.download(d_delta, like), info


@pytest.mark.parametrize("precision", ["double", "mixed", "float"])
def test_no_alloc_solve_matches_the_allocating_solve_twice(lib, hip, precision):
    # 6,000 nodes take the persistent grid kernel (47 groups of two node
    # slots) on a stream of the caller's, then 100 nodes the single-workgroup
    # kernel on the default stream, on the same array.
    stream = hip.stream()
    try:
        for n, seed, s in ((6000, 1, stream), (100, 2, None)):
            system = random_system(n, seed)
            expected, expected_info = allocating_solve(lib, *system, precision)
            got, info = noalloc_solve(lib, hip, *system, precision, s)
            assert expected_info[2] == 1.0 and info[2] == 1.0, (expected_info, info)
            assert np.array_equal(got, expected), (n, precision, np.abs(got - expected).max())
            assert np.array_equal(info, expected_info)
    finally:
        hip.close(stream)


def test_no_alloc_solve_rejects_sizes_beyond_its_array(lib, hip):
    small = hip.upload(np.zeros(64, dtype=np.int32))
    values = hip.upload(np.zeros(36 * 64))
    info = np.zeros(4)
    too_many = 16384 + 1
    status = lib.nricp_hip_solve_bsr_noalloc(small, small, small, too_many, too_many, values, values, DAMPING, 1, 0,
                                             1e-12, 100, values, ptr(info), None)
    assert status == 1
    assert b"node_count" in lib.nricp_hip_last_error()
    status = lib.nricp_hip_solve_bsr_noalloc(small, small, small, 8, 262144 + 1, values, values, DAMPING, 1, 1, 1e-12,
                                             100, values, ptr(info), None)
    assert status == 1
    assert b"block_count" in lib.nricp_hip_last_error()
