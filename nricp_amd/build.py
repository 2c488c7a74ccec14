"""Build recipe of the device library and the Python module, in tree.

* ``libnricp_hip.so`` -- every ``csrc/*.hip`` compiled by hipcc for gfx950
  into one shared library with the C ABI of ``include/nricp_hip.h``.
* ``nricp_dp<EXT_SUFFIX>`` -- the pybind11 module (``csrc/module.cpp``),
  linked against the library with an ``$ORIGIN`` run path.

Both land in this package directory. Kernels are compiled with
``-ffp-contract=off`` (the FP64 paths are compared bit for bit with the CPU
route) and with packed FP32 instructions disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
INCLUDE = PACKAGE.parent / "include"
BUILD = PACKAGE / "_build"
ARCH = os.environ.get("NRICP_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def library_path() -> Path:
    return PACKAGE / "libnricp_hip.so"


def module_path() -> Path:
    return PACKAGE / ("nricp_dp" + sysconfig.get_config_var("EXT_SUFFIX"))


def _run(cmd: list[str]) -> None:
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + result.stdout + result.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    if result.stderr.strip():
        # keep real warnings visible, drop the known host-pass feature note
        lines = [l for l in result.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


def _hip_version() -> str:
    try:
        out = subprocess.run(["hipconfig", "--version"], capture_output=True, text=True).stdout.strip()
        return out or "unknown"
    except OSError:
        return "unknown"


def build_library(jobs: int | None = None) -> Path:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    version = _hip_version()
    objects = []

    def compile_one(src: Path) -> Path:
        obj = BUILD / (src.stem + ".o")
        _run([hipcc, *HIP_FLAGS, f'-DNRICP_HIP_RUNTIME_VERSION="{version}"', f"-I{INCLUDE}", "-c", str(src),
              "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 4)) as pool:
        objects = list(pool.map(compile_one, sources))
    out = library_path()
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(out) + ".tmp"])
    os.replace(str(out) + ".tmp", out)
    return out


def build_module() -> Path:
    import pybind11

    out = module_path()
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
          f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}", f"-I{INCLUDE}",
          str(CSRC / "module.cpp"), f"-L{PACKAGE}", "-lnricp_hip", "-Wl,-rpath,$ORIGIN",
          "-o", str(out) + ".tmp"])
    os.replace(str(out) + ".tmp", out)
    return out


def build(jobs: int | None = None) -> None:
    build_library(jobs)
    build_module()


if __name__ == "__main__":
    build()
