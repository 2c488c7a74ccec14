"""The device's stable LSD radix sort (nricp_amd/csrc/primitives.hip), which
the hierarchy build, the node grid, the level layouts and the distance
quantiles rest on, against std::stable_sort: bit-identical keys and carried
values over tile-boundary sizes, 1 to 8 digit passes, long runs of equal keys
and many sorts in a row on one scratch. The program is built by
tests/native/build_native.py (part of the package build)."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent / "native"))
import build_native  # noqa: E402


@pytest.mark.gpu
def test_radix_sort_matches_stable_sort():
    exe = build_native.sort_test_path()
    if not exe.exists():
        pytest.fail(f"{exe} is not built: run `python __graft_entry__.py` first")
    print("sort_test: running", flush=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(out.stdout[-3000:], flush=True)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    assert out.stdout.strip().endswith("ALL OK")
