"""Build recipe for the CPU oracle (test infrastructure only).

* ``build_oracle()`` compiles this directory's FP64 restatement into the
  Python module ``oracle/nricp_oracle*.so``.
* ``build_ref()`` compiles the reference's own Eigen-free translation units
  (read where they lie under /root/reference/src) together with
  ``oracle/ref_shim/ref_capi.cpp`` -- a C entry layer written for this checker,
  not a stand-in for anything the reference lacks -- into
  ``oracle/_ref/libnricp_ref.so``.  Reference functions stay hidden and the
  linker drops what nothing calls, so the units that need Eigen
  (core/deformation.cpp, solver/direct_solve.cpp) are simply not part of it.

Both use ``-ffp-contract=off`` so that no fused multiply-add changes a rounding
relative to the other builds that are compared bit for bit.
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from pathlib import Path

HERE = Path(__file__).resolve().parent
REFERENCE_SRC = Path("/root/reference/src")
REF_UNITS = [
    "core/mesh.cpp",
    "core/sampling.cpp",
    "core/graph.cpp",
    "core/constraints.cpp",
    "core/report.cpp",
    "geometry/closest_point.cpp",
    "energy/robust.cpp",
    "solver/correspondence.cpp",
    "solver/normal_equations.cpp",
]
FLAGS = ["-O2", "-std=c++20", "-fPIC", "-ffp-contract=off"]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def oracle_module_path() -> Path:
    return HERE / ("nricp_oracle" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_oracle() -> Path:
    import pybind11

    out = oracle_module_path()
    sources = [HERE / "src" / n for n in ("oracle_core.cpp", "oracle_solve.cpp", "oracle_module.cpp")]
    cmd = ["g++", *FLAGS, "-shared", "-fvisibility=hidden",
           f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
           *map(str, sources), "-o", str(out) + ".tmp"]
    _run(cmd)
    os.replace(str(out) + ".tmp", out)
    return out


def ref_library_path() -> Path:
    return HERE / "_ref" / "libnricp_ref.so"


def build_ref() -> Path | None:
    """Compile the reference's Eigen-free units; returns None when absent."""
    if not REFERENCE_SRC.is_dir():
        return None
    out = ref_library_path()
    out.parent.mkdir(parents=True, exist_ok=True)
    sources = [str(REFERENCE_SRC / u) for u in REF_UNITS]
    sources.append(str(HERE / "ref_shim" / "ref_capi.cpp"))
    cmd = ["g++", *FLAGS, "-shared", "-fvisibility=hidden", "-ffunction-sections",
           "-fdata-sections", "-Wl,--gc-sections", "-Wl,--no-undefined",
           f"-I{REFERENCE_SRC}", *sources, "-o", str(out) + ".tmp"]
    _run(cmd)
    os.replace(str(out) + ".tmp", out)
    return out


if __name__ == "__main__":
    build_oracle()
    if "--no-ref" not in sys.argv:
        build_ref()
