"""The reference's recorded benchmark numbers, reproduced on the GPU.

pycfd pins its benchmark results in ``pycfd/tests/baselines.json`` and checks
them with ``test_regression.py`` (relative tolerance 1e-4, L-infinity metrics
3e-3, structurally zero quantities against absolute bounds).  The same
numbers, copied from baselines.json (fast tier, lines 52-88; full tier, lines
133-148), are checked here against the HIP path running the same cases: the
lid-driven cavity to steady state and the Re = 20 cylinder to t = 20.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1.0e-4       # test_regression.py:40
LINF_RTOL = 3.0e-3  # test_regression.py:78
# The FP64 GPU path reproduces these numbers far inside the reference's own
# cross-platform tolerances; the tight bound keeps it that way.
TIGHT = 1.0e-9


def _cavity(n: int, re: float = 100.0):
    """lid_driven_cavity.build: unit square, lid u = 1, dt ceiling 0.01, CFL
    0.4, steady_tol 1e-6, t_end max(25, 0.1 Re + 15), the direct solver."""
    from pycfd_amd import cases
    from pycfd_amd.physics.incompressible import Simulation
    cfg, _, _, _ = cases.cavity(nx=n, re=re, pressure_solver="direct", cfl_max=0.4,
                                dt=0.01, steady_tol=1.0e-6,
                                t_end=float(max(25.0, 0.1 * re + 15.0)))
    return Simulation(cfg)


def test_cavity_re100_48_matches_baseline(gpu_available):
    """baselines.json fast tier, cavity_re100 (48 x 48)."""
    from pycfd_amd.analysis.validation import ghia_errors
    sim = _cavity(48)
    res = sim.run()
    m = ghia_errors(sim.fields, 100)
    print(res.steps, res.time, m)
    assert res.converged
    assert res.steps == 3847
    assert res.time == pytest.approx(22.093407698376012, rel=RTOL)
    assert m["ghia_u_L2"] == pytest.approx(0.01698814301051285, rel=TIGHT)
    assert m["ghia_v_L2"] == pytest.approx(0.008059260315997074, rel=TIGHT)
    assert m["ghia_u_Linf"] == pytest.approx(0.06935902762975821, rel=LINF_RTOL)
    assert m["ghia_v_Linf"] == pytest.approx(0.0204928114358639, rel=LINF_RTOL)
    assert sim.solver.max_divergence(sim.fields) < 1e-10


def test_cavity_re100_128_ghia(gpu_available):
    """baselines.json full tier, cavity_re100_full (128 x 128, 18085 steps to
    steady state): centre-line L2 0.0067 / 0.0053, well inside 0.015."""
    from pycfd_amd.analysis.validation import ghia_errors
    sim = _cavity(128)
    res = sim.run()
    m = ghia_errors(sim.fields, 100)
    print(res.steps, res.time, res.wall_time, m)
    assert res.converged
    assert res.steps == 18085
    assert m["ghia_u_L2"] == pytest.approx(0.006666347306871988, rel=TIGHT)
    assert m["ghia_v_L2"] == pytest.approx(0.005332912643236642, rel=TIGHT)
    assert m["ghia_u_Linf"] == pytest.approx(0.025689443483192687, rel=LINF_RTOL)
    assert m["ghia_v_Linf"] == pytest.approx(0.009120649538648329, rel=LINF_RTOL)
    assert max(m["ghia_u_L2"], m["ghia_v_L2"]) <= 0.015
    assert sim.solver.max_divergence(sim.fields) < 1e-12


def test_cylinder_re20_matches_baseline(gpu_available):
    """baselines.json fast tier, cylinder_re20: cylinder_flow.build defaults
    (16 x 8 D domain, D = 1 at x = 4, uniform inflow, pressure outlet, slip
    walls, no perturbation below Re 50), 128 x 64, t_end = 20; Cd averaged over
    the second half of a record taken every 10 steps (cylinder_flow.py:458-492)."""
    from pycfd_amd import cases
    from pycfd_amd.analysis.postprocess import force_coefficients, late_time_mean
    from pycfd_amd.physics.incompressible import Simulation
    cfg, body, u0, _ = cases.cylinder(nx=128, ny=64, re=20.0, domain_length=16.0,
                                      domain_height=8.0, pressure_solver="direct",
                                      outlet="pressure_outlet", t_end=20.0)
    sim = Simulation(cfg, obstacle=body, u_init=u0)
    t, cd, cl = [], [], []

    def record(fields, info):
        f = force_coefficients(sim.solver.body_force_reaction, 1.0, cfg.l_ref)
        t.append(fields.t)
        cd.append(f[0])
        cl.append(f[1])

    res = sim.run(callback=record, callback_every=10)
    t, cd, cl = map(np.asarray, (t, cd, cl))
    settled = t > 0.5 * t[-1]
    cd_mean = late_time_mean(t, cd, 0.5)
    cl_rms = float(np.sqrt(np.mean(cl[settled] ** 2)))
    print(res.steps, cd_mean, cl_rms)
    assert res.steps == 1001
    assert res.time == pytest.approx(20.0, rel=RTOL)
    assert cd_mean == pytest.approx(2.471423208774098, rel=TIGHT)
    assert cl_rms < 1e-8
    assert body.characteristic_length == pytest.approx(1.0)
    assert sim.solver.max_divergence(sim.fields) < 1e-10
