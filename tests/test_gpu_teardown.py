"""Process exit with device contexts left open.

Contexts are closed by the library's exit hook (registered after torch has
initialised the device), and a context's finaliser does nothing while the
interpreter is finalising, so no device call runs after the HIP runtime or a
profiler's tool library has shut down.  The child program leaves two contexts
open, one with the live kernel clock on and the cooperative chain mode asked
for (kept when the device accepts the launch; it may refuse one while this test
runner holds the device).
"""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

CHILD = Path(__file__).resolve().parent / "teardown_child.py"


def test_exit_with_open_contexts_is_clean(gpu_available):
    r = subprocess.run([sys.executable, str(CHILD)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "exiting with contexts open" in r.stdout
