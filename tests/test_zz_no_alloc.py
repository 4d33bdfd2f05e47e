"""The allocation-free diagnostics entry point against the allocating one.

pcfd_diagnostics_noalloc keeps the reduction's partial sums in a fixed
device array instead of the context's buffers and writes to a device output;
it must return the same bits as pcfd_diagnostics, call after call, on grids
of different sizes sharing that array, and refuse a grid too large for it.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.cases_def import initial, oracle_case, product_config

pytestmark = pytest.mark.gpu


def _sim(name, steps, **kw):
    from pycfd_amd.physics.incompressible import Simulation
    cfg, body = product_config(name, **kw)
    u0, v0 = initial(name, oracle_case(name))
    sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
    sim.run(max_steps=steps)
    return sim


def _both(sim, previous):
    """(allocating, allocation-free) results for the same inputs through the C-ABI."""
    import torch
    from pycfd_amd import _lib
    lib = _lib.load()
    ctx, f = sim.solver.ctx, sim.fields
    ptr = lambda t: C.c_void_p(0 if t is None else t.data_ptr())  # noqa: E731
    up = previous.u_dev if previous is not None else None
    vp = previous.v_dev if previous is not None else None
    host = np.zeros(16)
    _lib.check(lib.pcfd_diagnostics(ctx.handle, ptr(f.u_dev), ptr(f.v_dev), ptr(f.p_dev),
                                    ptr(up), ptr(vp), host.ctypes.data_as(C.POINTER(C.c_double))),
               "pcfd_diagnostics")
    out = torch.full((9,), float("nan"), dtype=torch.float64, device="cuda")
    _lib.check(lib.pcfd_diagnostics_noalloc(ctx.handle, ptr(f.u_dev), ptr(f.v_dev),
                                            ptr(f.p_dev), ptr(up), ptr(vp), ptr(out)),
               "pcfd_diagnostics_noalloc")
    torch.cuda.synchronize()
    return host[:9], out.cpu().numpy()


def test_noalloc_diagnostics_equal_the_allocating_entry_point(gpu_available):
    """Two calls in a row, the second on a smaller grid: bit-identical to
    pcfd_diagnostics (same kernels and the same fold order)."""
    import torch
    from pycfd_amd.core.fields import FlowField
    big = _sim("cyl_outlet", 3, pressure_solver="mgcg")
    small = _sim("cavity", 2, pressure_solver="mgcg")
    assert big.fields.u_dev.numel() > small.fields.u_dev.numel()
    for sim in (big, small):
        prev = FlowField(sim.mesh, sim.fields.u_dev * 0.5, sim.fields.v_dev + 0.25,
                         torch.zeros_like(sim.fields.p_dev))
        for previous in (prev, None):
            a, b = _both(sim, previous)
            # without a previous state the max|du| slots (3, 4) carry no data
            data = [q for q in range(9) if previous is not None or q not in (3, 4)]
            assert np.all(np.isfinite(b[data]))
            assert np.array_equal(a, b), (a, b)
        assert np.array_equal(_both(sim, prev)[1][7:9], sim.solver.ctx.force()[:2])
    # and once more on the first grid, after the smaller one used the array
    a, b = _both(big, None)
    assert np.array_equal(a, b)


def test_noalloc_diagnostics_refuses_a_grid_beyond_its_scratch(gpu_available):
    """A 4600 x 4600 grid needs 20718 workgroups x 7 partials (1.16 MB),
    more than the 1 MiB array: refused with a clear error before any launch."""
    import torch
    from pycfd_amd import _lib
    from pycfd_amd.core.engine import DeviceContext
    ctx = DeviceContext(4600, 4600, 1.0 / 4600, 1.0 / 4600, 0.01)
    try:
        dummy = torch.zeros(16, dtype=torch.float64, device="cuda")
        p = C.c_void_p(dummy.data_ptr())
        rc = _lib.load().pcfd_diagnostics_noalloc(ctx.handle, p, p, None, None, None, p)
        assert rc == _lib.PCFD_ERR_BAD_ARG
        assert "maximum" in _lib.last_error()
    finally:
        ctx.close()
