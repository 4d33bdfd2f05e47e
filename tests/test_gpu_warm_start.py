"""The warm start by extrapolation in time (ProjectionSolver.warm_start_order).

Each RK substep's solve may start from the polynomial extrapolation of that
substep's last solutions instead of the last pressure solved (the reference's
``_p_guess``, order 0).  The solve stops on the same tolerance, so the fields
agree with the reference-started run (and the oracle) to the solver's
tolerance while later solves take fewer iterations; the ring of stored
solutions lives on the device, so captured steps replay it bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases_def import initial, oracle_case, product_config

pytestmark = pytest.mark.gpu

GRID = dict(nx=192, ny=96)


def _sim(order, graphs=False, scheme=None, name="cyl_outlet", solver="mgcg", tol=1e-12):
    from pycfd_amd.physics.incompressible import Simulation
    kw = dict(pressure_solver=solver, poisson_tol=tol, **GRID)
    if scheme:
        kw["time_scheme"] = scheme
    cfg, body = product_config(name, **kw)
    u0, v0 = initial(name, oracle_case(name, **GRID))
    sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
    sim.solver.use_graphs = graphs
    sim.solver.warm_start_order = order
    return sim


def _iterations(sim, steps):
    its = []
    for _ in range(steps):
        sim.run(max_steps=1)
        its.append(sim.solver.pressure_solver.last_iterations)
    return its


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_extrapolated_start_keeps_the_solution(gpu_available, order):
    steps = 14
    ref = _sim(0)
    ext = _sim(order)
    it_ref = _iterations(ref, steps)
    it_ext = _iterations(ext, steps)
    print(f"order {order}: iterations {it_ext} against {it_ref}")
    scale = max(1.0, float(np.abs(ref.fields.u).max()))
    assert np.abs(ext.fields.u - ref.fields.u).max() <= 1e-10 * scale
    assert np.abs(ext.fields.v - ref.fields.v).max() <= 1e-10 * scale
    assert np.abs(ext.fields.p - ref.fields.p).max() <= 1e-8 * max(1.0, np.abs(ref.fields.p).max())
    # the first steps fill the history; then no solve needs more iterations
    # than the reference-started one, and the sum drops
    late = slice(order + 2, None)
    assert all(a <= b for a, b in zip(it_ext[late], it_ref[late])), (it_ext, it_ref)
    assert sum(it_ext[late]) < sum(it_ref[late]), (it_ext, it_ref)


@pytest.mark.parametrize("solver", ["cg", "sor", "multigrid", "direct"])
def test_extrapolated_start_other_solvers(gpu_available, solver):
    """Every solver starts from the engine's guess, so the extrapolation
    applies to all of them: the fields stay those of the reference-started
    run to the solver's tolerance and the solves take no more iterations in
    all (SOR tests its residual every fifth sweep)."""
    steps = 10
    ref = _sim(0, solver=solver, tol=1e-11)
    ext = _sim(3, solver=solver, tol=1e-11)
    it_ref = _iterations(ref, steps)
    it_ext = _iterations(ext, steps)
    print(f"{solver}: iterations {it_ext} against {it_ref}")
    scale = max(1.0, float(np.abs(ref.fields.u).max()))
    assert np.abs(ext.fields.u - ref.fields.u).max() <= 1e-7 * scale
    assert np.abs(ext.fields.v - ref.fields.v).max() <= 1e-7 * scale
    late = slice(5, None)
    assert sum(it_ext[late]) <= sum(it_ref[late]), (it_ext, it_ref)


def test_extrapolated_start_matches_oracle(gpu_available):
    """Against the oracle's run (which starts every solve the reference's
    way): velocity to 1e-10 relative after 10 steps at tolerance 1e-12."""
    steps = 10
    sim = _sim(3)
    sim.run(max_steps=steps)
    case = oracle_case("cyl_outlet", pressure_solver="mgcg", tol=1e-12, **GRID)
    u0, v0 = initial("cyl_outlet", case)
    u, v, p, hist, t = O.run(case, steps, u0, v0)
    scale = max(1.0, float(np.abs(u).max()))
    assert np.abs(sim.fields.u - u).max() <= 1e-10 * scale
    assert np.abs(sim.fields.v - v).max() <= 1e-10 * scale


@pytest.mark.parametrize("scheme,order", [("rk3", 3), ("rk2", 3), ("rk3", 4)])
def test_extrapolated_start_graphs_equal_eager(gpu_available, scheme, order):
    """Captured steps keep the history's state on the device: replays give the
    eager run's fields and iteration counts bit for bit."""
    steps = 12
    a = _sim(order, graphs=False, scheme=scheme)
    b = _sim(order, graphs=True, scheme=scheme)
    ia = _iterations(a, steps)
    ib = _iterations(b, steps)
    assert b.solver.ctx.step_graph_stats()["launches"] >= steps - 2
    assert ia == ib
    assert np.array_equal(a.fields.u, b.fields.u)
    assert np.array_equal(a.fields.v, b.fields.v)
    assert np.array_equal(a.fields.p, b.fields.p)


def test_setting_the_guess_restarts_the_history(gpu_available):
    """A pressure guess set from outside starts the history afresh: the next
    steps are those of a run that enabled the extrapolation at that point."""
    steps = 6
    a = _sim(3)
    a.run(max_steps=steps)
    state = (a.fields.u.copy(), a.fields.v.copy(), a.solver.pressure_guess())
    a.solver.set_pressure_guess(state[2])
    a.run(max_steps=steps)
    b = _sim(0)
    b.run(max_steps=steps)
    b.solver.warm_start_order = 3
    b.solver.set_pressure_guess(b.solver.pressure_guess())
    b.run(max_steps=steps)
    # the first halves differ only in the solves' starts; the second halves
    # both restart the history, so they agree to the solver's tolerance
    scale = max(1.0, float(np.abs(a.fields.u).max()))
    assert np.abs(a.fields.u - b.fields.u).max() <= 1e-10 * scale


def test_order_is_checked(gpu_available):
    sim = _sim(0)
    with pytest.raises(ValueError, match="0-4"):
        sim.solver.warm_start_order = 5
    assert sim.solver.warm_start_order == 0
