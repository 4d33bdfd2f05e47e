"""Jacobi-CG on small grids as one workgroup (k_jpcg_small) against the
two-launch iteration (k_jpcg_dir / k_jpcg_update_f, each folding the other's
partial sums) and the oracle.

Grids of at most 2048 cells run the whole CG solve in one launch on one CU;
PCFD_SMALL_CG=0 selects the multi-launch iteration on the same grid.  Both run
scipy's CG recurrence (pressure.py:366-398) and differ only in the order the
inner products are summed, so they stop within an iteration or two of each
other and agree to the solver tolerance.  Covered: every boundary family of
the small cases (walls, periodic, inlet/outlet with a body, pressure outlet),
stretched meshes, both register layouts (1 and 2 cells per thread, the second
up to the limit), and grids past the limit (the two-launch iteration on both
settings).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases_def import initial, oracle_case, product_config

pytestmark = pytest.mark.gpu


def _run(name, steps, small, monkeypatch, **kw):
    from pycfd_amd.physics.incompressible import Simulation
    monkeypatch.setenv("PCFD_SMALL_CG", "1" if small else "0")
    cfg, body = product_config(name, pressure_solver="cg", poisson_tol=1e-12, **kw)
    case = oracle_case(name, **{k: v for k, v in kw.items() if k in ("nx", "ny")})
    u0, v0 = initial(name, case)
    sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
    res = sim.run(max_steps=steps)
    return sim, res


@pytest.mark.parametrize("name,kw", [
    ("cavity", {}),                                   # 480 cells: 1 per thread
    ("periodic", {}),                                 # periodic both ways
    ("cyl_outlet", {}),                               # inlet, outlet, body
    ("cyl_pout", {}),                                 # pressure outlet (non-singular)
    ("s_cavity", {}),                                 # stretched
    ("s_cyl", {}),                                    # stretched with a body
    ("channel", {"nx": 40, "ny": 36}),                # 1440 cells: 2 per thread, ragged
    ("cavity", {"nx": 64, "ny": 32}),                 # 2048 cells: the limit
    ("cavity", {"nx": 66, "ny": 32}),                 # just past it
    ("channel", {"nx": 96, "ny": 40}),                # 3840 cells
])
def test_small_cg_matches_multi_launch(gpu_available, monkeypatch, name, kw):
    steps = 5
    sa, ra = _run(name, steps, True, monkeypatch, **kw)
    sb, rb = _run(name, steps, False, monkeypatch, **kw)
    for a, b in ((sa.fields.u, sb.fields.u), (sa.fields.v, sb.fields.v)):
        scale = max(1.0, float(np.abs(b).max()))
        assert np.abs(a - b).max() <= 1e-10 * scale
    np.testing.assert_allclose(ra.history["dt"], rb.history["dt"], rtol=1e-12)
    ia = sa.solver.pressure_solver.last_iterations
    ib = sb.solver.pressure_solver.last_iterations
    assert abs(ia - ib) <= 2, (ia, ib)
    assert max(ra.history["max_div"]) <= 1e-9
    assert sa.solver.pressure_solver.last_residual <= 1e-11


def test_small_cg_matches_oracle(gpu_available, monkeypatch):
    """A 45 x 30 cavity (ragged last cell block) against the oracle's CG run."""
    name, kw, steps = "cavity", {"nx": 45, "ny": 30}, 4
    sim, res = _run(name, steps, True, monkeypatch, **kw)
    case = oracle_case(name, pressure_solver="cg", tol=1e-12, **kw)
    u, v, p, hist, t = O.run(case, steps)
    assert np.abs(sim.fields.u - u).max() <= 1e-10
    assert np.abs(sim.fields.v - v).max() <= 1e-10
    np.testing.assert_allclose(res.history["dt"], hist["dt"], rtol=1e-12)


def test_small_cg_respects_maxiter(gpu_available, monkeypatch):
    """A solve capped below what it needs stops at exactly maxiter, unconverged,
    on both paths (the reference raises; the error text names the residual)."""
    from pycfd_amd.core.pressure import assemble_poisson_matrix, make_pressure_solver
    from pycfd_amd.core.mesh import StructuredMesh
    mesh = StructuredMesh(48, 40, 1.0, 1.0)
    system = assemble_poisson_matrix(mesh)
    rng = np.random.default_rng(7)
    rhs = rng.standard_normal((48, 40))
    out = {}
    for small in (True, False):
        monkeypatch.setenv("PCFD_SMALL_CG", "1" if small else "0")
        solver = make_pressure_solver("cg", system, tol=1e-14, maxiter=7)
        with pytest.raises(RuntimeError, match="failed to converge"):
            solver.solve(rhs)
        out[small] = solver.last_iterations
    assert out[True] == out[False] == 7
