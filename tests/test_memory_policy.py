"""The device-memory cache's ordering policy, on the host: a released block
is handed back only to requests ordered on its own stream until that stream
retires, idle blocks go to anyone, devices and size classes never mix
(nricp_amd/csrc/stream_pool.hpp, the policy memory.hip runs on the device).
The C++ test is compiled with g++ and run; no GPU is involved."""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_stream_ordered_pool_rules(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.fail("g++ is needed to build the host test")
    exe = tmp_path / "stream_pool_test"
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", f"-I{ROOT / 'nricp_amd' / 'csrc'}",
                    str(ROOT / "tests" / "native" / "stream_pool_test.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"


def test_xcd_work_mapping(tmp_path):
    """The gather kernels' workgroup -> work mapping (nricp_amd/csrc/xcd_map.hpp):
    a bijection that gives each XCD one contiguous run of each kind of work."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.fail("g++ is needed to build the host test")
    exe = tmp_path / "xcd_map_test"
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", f"-I{ROOT / 'nricp_amd' / 'csrc'}",
                    str(ROOT / "tests" / "native" / "xcd_map_test.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
