"""Builds the native device tests in tree (``tests/native/_build``): host C++
programs that drive the library's internal primitives directly, linked
against ``nricp_amd/libnricp_hip.so``. Compiled here on the CPU, run by the
``gpu``-marked tests."""

from __future__ import annotations

import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
OUT = HERE / "_build"
ARCH = "gfx950"


def sort_test_path() -> Path:
    return OUT / "sort_test"


def build() -> Path:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    OUT.mkdir(exist_ok=True)
    exe = sort_test_path()
    cmd = [hipcc, f"--offload-arch={ARCH}", "-O2", "-std=c++17",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
           f"-I{ROOT / 'nricp_amd' / 'csrc'}", f"-I{ROOT / 'include'}", str(HERE / "sort_test.cpp"),
           f"-L{ROOT / 'nricp_amd'}", "-lnricp_hip", "-Wl,-rpath,$ORIGIN/../../../nricp_amd",
           "-o", str(exe) + ".tmp"]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + result.stdout + result.stderr)
        raise RuntimeError("building tests/native/sort_test.cpp failed")
    Path(str(exe) + ".tmp").replace(exe)
    return exe


if __name__ == "__main__":
    print(build())
