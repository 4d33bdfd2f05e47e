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


def test_taylor_green_convergence_matches_baseline(gpu_available):
    """baselines.json fast tier, taylor_green_convergence: taylor_green.build
    (nu = 0.05, t_end = 0.5, dt = 0.4 * 2 pi / N, fixed dt, direct solver) at
    N = 16, 32, 64, 128; errors of the face velocities against the exact vortex
    (taylor_green.py:118-132) and the observed orders."""
    from pycfd_amd.analysis.validation import convergence_orders, taylor_green
    from pycfd_amd.config import BCKind, BCSpec, SimulationConfig
    from pycfd_amd.core.mesh import StructuredMesh
    from pycfd_amd.physics.incompressible import Simulation
    nu, L = 0.05, 2.0 * np.pi
    ns, l2, linf = (16, 32, 64, 128), [], []
    for n in ns:
        cfg = SimulationConfig(nx=n, ny=n, lx=L, ly=L, re=1.0 / nu, u_ref=1.0, l_ref=1.0,
                               dt=0.4 * L / n, t_end=0.5, adaptive_dt=False,
                               boundary_config={w: BCSpec(BCKind.PERIODIC)
                                                for w in ("left", "right", "bottom", "top")})
        mesh = StructuredMesh.from_config(cfg)
        Xu, Yu = mesh.u_grid()
        Xv, Yv = mesh.v_grid()
        sim = Simulation(cfg, u_init=taylor_green(Xu, Yu, 0.0, nu, "u"),
                         v_init=taylor_green(Xv, Yv, 0.0, nu, "v"))
        sim.run()
        t = sim.fields.t
        du = sim.fields.u_phys[:-1, :] - taylor_green(Xu, Yu, t, nu, "u")[:-1, :]
        dv = sim.fields.v_phys[:, :-1] - taylor_green(Xv, Yv, t, nu, "v")[:, :-1]
        l2.append(float(np.sqrt((np.sum(du ** 2) + np.sum(dv ** 2)) / (du.size + dv.size))))
        linf.append(float(max(np.abs(du).max(), np.abs(dv).max())))
        assert sim.solver.max_divergence(sim.fields) < 1e-10
    o2, oi = convergence_orders(ns, l2)[-1], convergence_orders(ns, linf)[-1]
    print(l2, linf, o2, oi)
    assert o2 == pytest.approx(1.999669973940691, rel=RTOL)
    assert oi == pytest.approx(1.9965084715978074, rel=LINF_RTOL)
    assert l2[-1] == pytest.approx(4.77475461891742e-06, rel=RTOL)
    assert linf[-1] == pytest.approx(9.558930812225697e-06, rel=LINF_RTOL)


@pytest.mark.parametrize("mode", ["periodic", "developing"])
def test_channel_matches_baseline(gpu_available, mode):
    """baselines.json fast tier, channel_periodic (16 x 32, body force) and
    channel_developing (48 x 24, uniform inlet, zero-gradient outlet):
    channel_flow.build at Re 10 to steady_tol 1e-9, profile against the
    parabola of equal flux (channel_flow.py:162-208)."""
    from pycfd_amd.analysis.validation import l2_error, poiseuille_profile
    from pycfd_amd.config import BCKind, BCSpec, SimulationConfig
    from pycfd_amd.physics.incompressible import Simulation
    re, nu = 10.0, 0.1
    if mode == "periodic":
        nx, ny, length, force, u0 = 16, 32, 2.0, (8.0 * nu, 0.0), 0.0
        bcs = {"left": BCSpec(BCKind.PERIODIC), "right": BCSpec(BCKind.PERIODIC),
               "bottom": BCSpec(BCKind.NO_SLIP), "top": BCSpec(BCKind.NO_SLIP)}
        ref = dict(centerline_error_pct=1.0090861479739033e-07,
                   profile_L2_relative=0.0009104384738361293,
                   u_max_numerical=0.9999999989909139, mean_velocity=0.6679687493565628,
                   steps=5722, final_time=21.036764705883428)
    else:
        nx, ny, length, force, u0 = 48, 24, 6.0, (0.0, 0.0), 1.0
        bcs = {"left": BCSpec(BCKind.INLET, velocity=1.0), "right": BCSpec(BCKind.OUTLET),
               "bottom": BCSpec(BCKind.NO_SLIP), "top": BCSpec(BCKind.NO_SLIP)}
        ref = dict(centerline_error_pct=0.002499413032156876,
                   profile_L2_relative=0.0023089094005366083,
                   u_max_numerical=1.4999625088045176, mean_velocity=0.9999999999999997,
                   steps=447, final_time=2.793750000000023)
    cfg = SimulationConfig(nx=nx, ny=ny, lx=length, ly=1.0, re=re, u_ref=1.0, l_ref=1.0,
                           dt=0.01, t_end=80.0, cfl_max=0.4, body_force=force,
                           boundary_config=bcs, steady_tol=1.0e-9)
    sim = Simulation(cfg, u_init=u0)
    res = sim.run()
    uc, _ = sim.fields.cell_velocities()
    profile = uc[nx // 2 if mode == "periodic" else nx - 1, :]
    mean = float(profile.mean())
    peak = 1.0 if mode == "periodic" else 1.5
    got = dict(centerline_error_pct=abs(profile.max() - peak) / peak * 100.0,
               profile_L2_relative=l2_error(profile, poiseuille_profile(sim.mesh.yc, 1.0,
                                                                        1.5 * mean),
                                            relative=True),
               u_max_numerical=float(profile.max()), mean_velocity=mean,
               steps=res.steps, final_time=res.time)
    print(mode, got)
    assert res.converged
    assert got["steps"] == ref["steps"]
    for k in ("profile_L2_relative", "u_max_numerical", "mean_velocity", "final_time"):
        assert got[k] == pytest.approx(ref[k], rel=RTOL), k
    assert got["centerline_error_pct"] == pytest.approx(ref["centerline_error_pct"],
                                                        rel=1e-3, abs=1e-9)
    assert sim.solver.max_divergence(sim.fields) < 1e-10
