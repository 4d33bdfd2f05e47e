"""Shared test configuration.

Markers:
  gpu  -- needs a HIP device and the built libpycfd_hip.so; every GPU test
          compares the HIP engine with the CPU oracle (oracle/oracle.py) or
          with the golden vectors recorded from the reference.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built engine")


def golden(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu_available():
    try:
        import torch
        ok = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        ok = False
    if not ok:
        pytest.skip("no HIP device")
    from pycfd_amd import _lib
    _lib.load()
    return True
