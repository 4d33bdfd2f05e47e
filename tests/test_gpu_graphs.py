"""Whole steps captured as HIP graphs against the same steps launched kernel
by kernel.

A captured step replays the eager step's kernels in the eager order, with
dt and the upwind blend written by its first node and a fixed iteration
budget per solve (iterations past convergence are no-op kernels), so every
replay must reproduce the eager path bit for bit: fields, pressure, forces
and the per-step diagnostics.  Replays alternate with the program's own
eager work (the per-step reductions, fresh output buffers), and a replay
whose budget runs out is redone eagerly.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests.cases_def import initial, oracle_case, product_config

pytestmark = pytest.mark.gpu

CASES = [("cavity", dict(pressure_solver="mgcg")),
         ("cavity", dict(pressure_solver="cg")),
         ("cavity", dict(pressure_solver="sor")),
         ("cavity", dict(pressure_solver="jacobi", poisson_tol=1e-8)),
         ("cyl_outlet", dict(pressure_solver="mgcg", precision="mixed")),
         ("cyl_pout", dict(pressure_solver="direct")),
         ("periodic", dict(pressure_solver="multigrid", mg_sweeps=2)),
         ("channel", dict(pressure_solver="mgcg", time_scheme="rk2")),
         ("cyl_outlet", dict(pressure_solver="mgcg", time_scheme="euler",
                             advection_scheme="upwind")),
         ("cavity", dict(pressure_solver="mgcg", use_les=True)),
         ("s_cyl", dict(pressure_solver="mgcg"))]


def _run(name, kw, graphs, steps):
    from pycfd_amd.physics.incompressible import Simulation
    cfg, body = product_config(name, **kw)
    u0, v0 = initial(name, oracle_case(name))
    sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
    sim.solver.use_graphs = graphs
    res = sim.run(max_steps=steps)
    return sim, res


@pytest.mark.parametrize("name,kw", CASES, ids=[f"{c[0]}-{'-'.join(map(str, c[1].values()))}"
                                               for c in CASES])
def test_graph_steps_equal_eager_steps(gpu_available, name, kw):
    steps = 12
    a, ra = _run(name, kw, False, steps)
    b, rb = _run(name, kw, True, steps)
    st = b.solver.ctx.step_graph_stats()
    assert st["launches"] >= steps - 2, st        # the first step measures the counts
    for k in ("dt", "max_div", "kinetic_energy", "fx", "fy", "residual"):
        assert np.array_equal(np.asarray(ra.history[k]), np.asarray(rb.history[k])), k
    assert np.array_equal(a.fields.u, b.fields.u)
    assert np.array_equal(a.fields.v, b.fields.v)
    assert np.array_equal(a.fields.p, b.fields.p)
    assert a.solver.pressure_solver.last_iterations == b.solver.pressure_solver.last_iterations
    assert a.solver.body_force_reaction == b.solver.body_force_reaction


def test_budget_shortfall_redoes_the_step_eagerly(gpu_available, monkeypatch):
    """A warm start replaced by large noise before a replayed step makes its
    first solve need far more iterations than the captured budget: the replay
    is redone eagerly from the same start, and the run stays equal to the
    eager run.  (The iterated Jacobi-CG: the one-workgroup solve of tiny
    grids runs to convergence inside its single launch and has no budget.)"""
    monkeypatch.setenv("PCFD_SMALL_CG", "0")
    results = []
    for graphs in (False, True):
        from pycfd_amd.physics.incompressible import Simulation
        # Jacobi-CG: at this size multigrid-CG is one exact dense solve
        cfg, body = product_config("cyl_outlet", pressure_solver="cg")
        u0, v0 = initial("cyl_outlet", oracle_case("cyl_outlet"))
        sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
        sim.solver.use_graphs = graphs
        sim.run(max_steps=5)
        noise = np.random.default_rng(3).standard_normal(sim.mesh.shape) * 1e3
        sim.solver.set_pressure_guess(noise)
        sim.run(max_steps=5)
        results.append((sim.fields.u.copy(), sim.fields.p.copy(),
                        sim.solver.ctx.step_graph_stats()))
    (ua, pa, _), (ub, pb, st) = results
    assert st["fallbacks"] >= 1 and st["launches"] >= 6, st
    assert np.array_equal(ua, ub) and np.array_equal(pa, pb)


def test_graphs_toggled_mid_run_match_eager(gpu_available):
    """Graph steps, then eager steps, then graph steps again (recaptured
    against whatever buffers the allocator hands out) equal a run that
    never used graphs."""
    from pycfd_amd.physics.incompressible import Simulation
    out = []
    for pattern in ((False, False, False), (True, False, True)):
        cfg, body = product_config("cyl_pout", pressure_solver="mgcg")
        u0, v0 = initial("cyl_pout", oracle_case("cyl_pout"))
        sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
        for g in pattern:
            sim.solver.use_graphs = g
            sim.run(max_steps=4)
        out.append((sim.fields.u.copy(), sim.fields.v.copy(), sim.fields.p.copy()))
    for x, y in zip(*out):
        assert np.array_equal(x, y)
