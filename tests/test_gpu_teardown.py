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


HOST_FIRST = """
import sys
sys.path.insert(0, {root!r})
from pycfd_amd.core.mesh import StructuredMesh
from pycfd_amd.core.pressure import assemble_poisson_matrix
from pycfd_amd.core.multigrid import MultigridHierarchy
S = assemble_poisson_matrix(StructuredMesh(64, 32, 2.0, 1.0))  # native host calls first
H = MultigridHierarchy(S.A, S.shape, S.fluid, S.singular)
from pycfd_amd.core.engine import DeviceContext
ctx = DeviceContext(64, 32, 2.0 / 64, 1.0 / 32, 0.01)
ctx.close()
print("context after host setup", flush=True)
"""


def test_device_after_native_host_setup(gpu_available):
    """The library's host-only entry points (assembly, hierarchy build) used
    before anything touches the device: the device is still found after (the
    library binds the HIP runtime torch ships, one per process)."""
    root = str(Path(__file__).resolve().parent.parent)
    r = subprocess.run([sys.executable, "-c", HOST_FIRST.format(root=root)],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "context after host setup" in r.stdout
