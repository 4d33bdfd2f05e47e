"""The in-project autotuner's smoother-sweep choice.

Tile shapes and kernel forms are measured at configure time (their choice
never changes results); the multigrid smoothing sweeps change the
preconditioner, so they are tuned only on request, by timing whole steps,
and every choice must still give the step to the solver tolerance.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases_def import initial, oracle_case, product_config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ps", ["mgcg", "multigrid"])
def test_tuned_sweeps_keep_the_step(gpu_available, ps):
    from pycfd_amd.physics.incompressible import Simulation
    name = "cyl_pout"
    grid = dict(nx=96, ny=48)
    cfg, body = product_config(name, pressure_solver=ps, poisson_tol=1e-12, **grid)
    case = oracle_case(name, pressure_solver="direct", tol=1e-12, **grid)
    u0, v0 = initial(name, case)
    sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
    table = sim.solver.tune_smoother_sweeps(sim.fields, 0.01, candidates=(1, 2, 3), steps=2)
    assert sorted(table) == [1, 2, 3]
    # the fewest sweeps within 3% of the fastest
    fastest = min(t["ms_per_step"] for t in table.values())
    best = min(k for k, t in table.items() if t["ms_per_step"] <= 1.03 * fastest)
    assert sim.solver.pressure_solver.smooth_sweeps == best
    assert all(t["iterations"] > 0 and t["ms_per_step"] > 0 for t in table.values())
    # more smoothing per cycle never needs more cycles on this case
    assert table[3]["iterations"] <= table[1]["iterations"]
    sim.run(max_steps=4)
    u, v, p, hist, t = O.run(case, 4, u0, v0)
    assert np.abs(sim.fields.u - u).max() <= 1e-10 * np.abs(u).max()
    assert np.abs(sim.fields.p - p).max() <= 1e-8 * max(np.abs(p).max(), 1.0)


def test_tuning_other_solvers_is_a_no_op(gpu_available):
    from pycfd_amd.physics.incompressible import Simulation
    cfg, body = product_config("cavity", pressure_solver="cg")
    sim = Simulation(cfg, obstacle=body)
    assert sim.solver.tune_smoother_sweeps(sim.fields, 0.01) == {}
