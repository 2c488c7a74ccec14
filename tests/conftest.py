"""Test configuration.

Two implementations of the same Python surface are exercised:

* ``oracle`` -- the FP64 CPU restatement of the reference's route A
  (``oracle/nricp_oracle*.so``), the checker;
* ``device`` -- ``nricp_dp``, the MI355X implementation under test.

Tests that take the ``impl`` fixture run once per implementation; the device
instance carries the ``gpu`` marker, so ``-m "not gpu"`` pins the oracle on the
CPU and ``-m gpu`` runs the same assertions on the device. A GPU test never
skips for a missing device or a missing extension: it fails.
"""

from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

# Read once, when HIP initialises (before any test touches the device). With
# the runtime's default of 4 hardware queues, concurrent jobs' streams share
# queues and their kernels serialise, which hides ordering mistakes between
# jobs; 16 lets the serving tests' jobs really overlap.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

ROOT = Path(__file__).resolve().parents[1]
for sub in ("", "nricp_amd", "oracle", "tests"):
    path = str(ROOT / sub) if sub else str(ROOT)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    import nricp_oracle

    return nricp_oracle


@pytest.fixture(scope="session")
def device():
    import nricp_dp

    info = nricp_dp.device_info()
    if not info["available"]:
        pytest.fail(f"the device module loaded but no HIP device is visible: {info['device']}")
    return nricp_dp


@pytest.fixture(scope="session")
def host_module():
    """nricp_dp for its host-side entry points (no device needed)."""
    import nricp_dp

    return nricp_dp


@pytest.fixture(params=["oracle", pytest.param("device", marks=pytest.mark.gpu)])
def impl(request):
    return request.getfixturevalue(request.param)


@pytest.fixture(params=["oracle", "host_module"])
def host_impl(request):
    """Both implementations of the host-only entry points (validation,
    cleaning, normalisation, report serialisation, block materialisation)."""
    return request.getfixturevalue(request.param)
