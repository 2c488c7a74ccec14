"""The C-ABI library: it loads without a GPU and exports every entry point
that include/nricp_hip.h declares; the status and error plumbing works
without touching a device."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "nricp_hip.h"
LIBRARY = ROOT / "nricp_amd" / "libnricp_hip.so"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = set(re.findall(r"\b(nricp_hip_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not LIBRARY.exists():
        pytest.fail(f"{LIBRARY} is not built (run __graft_entry__.build())")
    return ctypes.CDLL(str(LIBRARY))


def test_header_declares_the_session_and_operator_surface():
    names = declared()
    for required in ("nricp_hip_session_create", "nricp_hip_session_step", "nricp_hip_session_destroy",
                     "nricp_hip_closest_points", "nricp_hip_fused_system", "nricp_hip_solve_system",
                     "nricp_hip_last_error", "nricp_hip_device_info"):
        assert required in names
    assert len(names) >= 30


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_null_arguments_are_reported_not_crashed(lib):
    lib.nricp_hip_last_error.restype = ctypes.c_char_p
    status = lib.nricp_hip_session_step(None, None)
    assert status != 0
    assert b"null" in lib.nricp_hip_last_error()
