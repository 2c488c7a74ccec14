"""MI355X (gfx950) implementation of the NRICP inner loop.

The user-facing module is ``nricp_dp`` (built next to this file), with the
reference module's API; ``workloads`` generates the seeded benchmark jobs.
"""

from __future__ import annotations

import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)


def load():
    """Import the device module; fails loudly when it has not been built."""
    try:
        import nricp_dp  # noqa: PLC0415
    except ImportError as exc:  # pragma: no cover - depends on the build
        raise ImportError(
            "nricp_dp is not built: run `python __graft_entry__.py` (hipcc, gfx950) first"
        ) from exc
    return nricp_dp
