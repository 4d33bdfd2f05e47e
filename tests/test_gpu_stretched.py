"""Stretched meshes on the GPU (mesh.py:122-310 with the array-metric
operators of solver.py:334-683 and pressure.py:193-207).

Stencils and the wall passes bit for bit against vectors the reference
recorded (tests/golden/stretched.npz); the Poisson operator, V-cycle and whole
steps to the stated tolerances against the reference or the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases_def import STRETCHED, initial, oracle_case, product_config
from tests.conftest import golden

pytestmark = pytest.mark.gpu

NAMES = list(STRETCHED)


def _solver(name, **kw):
    from pycfd_amd.core.mesh import StructuredMesh
    from pycfd_amd.core.solver import ProjectionSolver
    cfg, body = product_config(name, **kw)
    mesh = StructuredMesh.from_config(cfg)
    return ProjectionSolver(cfg, mesh, body.mask if body else None)


def _sim(name, u0=None, v0=None, **kw):
    from pycfd_amd.physics.incompressible import Simulation
    cfg, body = product_config(name, **kw)
    return Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)


@pytest.mark.parametrize("name", NAMES)
def test_stretched_mesh_widths(gpu_available, name):
    """mesh.py:269-311 widths, bit for bit."""
    g = golden("stretched")
    s = _solver(name)
    assert np.array_equal(s.mesh.dx_cells, g[f"{name}_dx_cells"])
    assert np.array_equal(s.mesh.dy_cells, g[f"{name}_dy_cells"])


@pytest.mark.parametrize("name", NAMES)
def test_stretched_rhs_and_divergence_bitwise(gpu_available, name):
    """Weighted corner interpolation, per-face control volumes and the
    stretched viscous form (solver.py:334-465); divergence over each cell's
    own width (solver.py:532-543): bit for bit."""
    g = golden("stretched")
    s = _solver(name)
    u, v = g[f"{name}_u"], g[f"{name}_v"]
    for b in (0, 7):
        ru, rv = s.momentum_rhs(u, v, b / 10.0)
        assert np.array_equal(ru, g[f"{name}_b{b}_rhs_u"]), b
        assert np.array_equal(rv, g[f"{name}_b{b}_rhs_v"]), b
    assert np.array_equal(s.divergence(u, v), g[f"{name}_div"])


@pytest.mark.parametrize("name", NAMES)
def test_stretched_boundary_pass(gpu_available, name):
    """Predictor wall pass, width-weighted mass balance (boundary.py:438-487)
    and the immersed-boundary force over control-volume areas
    (solver.py:197-207): values to 1e-14, force to 1e-12 (tree sums against
    numpy's pairwise sums)."""
    from pycfd_amd.core.fields import to_device, to_host
    g = golden("stretched")
    s = _solver(name)
    ctx = s.ctx
    u, v = to_device(g[f"{name}_u"]), to_device(g[f"{name}_v"])
    ctx.apply_velocity_bc(u, v, True)
    ctx.mass_balance(u, v)
    ctx.mask_obstacle(u, v, 0.01, 1)
    np.testing.assert_allclose(to_host(u), g[f"{name}_pass_u"], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(to_host(v), g[f"{name}_pass_v"], rtol=1e-14, atol=1e-14)
    if s.has_obstacle:
        np.testing.assert_allclose(ctx.force()[:2], g[f"{name}_pass_force"], rtol=1e-12,
                                   atol=1e-12)


@pytest.mark.parametrize("name", NAMES + ["s_pchannel"])
def test_stretched_operator(gpu_available, name):
    """The area-scaled symmetric operator (pressure.py:193-207): the host
    assembly equals the reference's row scale and Dirichlet terms bit for bit,
    and the device's matrix-free product matches A x to 1e-14."""
    import torch
    s = _solver(name)
    A = s.poisson.A
    if name in STRETCHED:
        g = golden("stretched")
        x = g[f"{name}_Ax_x"]
        ref = g[f"{name}_Ax"]
        assert np.abs(A @ x - ref).max() <= 1e-14 * np.abs(ref).max()
        rs = np.broadcast_to(np.asarray(s.poisson.row_scale, dtype=float), x.shape)
        assert np.array_equal(rs, np.broadcast_to(g[f"{name}_row_scale"], x.shape))
        assert np.array_equal(s.poisson.bc_rhs, g[f"{name}_bc_rhs"])
    else:
        x = np.random.default_rng(3).standard_normal(A.shape[0])
        ref = A @ x
    assert abs(A - A.T).max() <= 1e-12 * abs(A).max()
    y = torch.empty(A.shape[0], dtype=torch.float64, device="cuda")
    s.ctx.poisson_apply(torch.as_tensor(x).cuda(), y)
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-14 * np.abs(ref).max()


CYCLE_MODES = {"fused": (True, -1, False), "passes": (False, -1, False),
               "wide": (True, -1, True), "chain1-wide": (True, 1, True)}


@pytest.mark.parametrize("name", NAMES + ["s_pchannel"])
@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("mode", list(CYCLE_MODES))
def test_stretched_vcycle_matches_oracle(gpu_available, name, sweeps, mode):
    """One V-cycle on a stretched operator (reference coarsest level, 96
    unknowns) vs the oracle's: FP64 to 1e-10, the FP32 cycle to 1e-4 of
    max|z|.  Modes: fused level-0 kernels with fused or pass-by-pass stored
    levels, the stored levels handed over with 32-bit class indices (what a
    stretched level past 65535 distinct rows uses), and that with the
    cooperative chain."""
    import torch
    fused, chain, wide = CYCLE_MODES[mode]
    case = oracle_case(name)
    S = O.assemble(case)
    H = O.Hierarchy(S.A, (case.nx, case.ny), S.fluid, S.singular, sweeps, coarse_max=96)
    r = np.random.default_rng(8).standard_normal(S.A.shape[0])
    ref = H.apply(r)
    for precision, tol in (("fp64", 1e-10), ("mixed", 1e-4)):
        s = _solver(name, pressure_solver="mgcg", mg_sweeps=sweeps, mg_coarse_max=96,
                    precision=precision)
        ps = s.pressure_solver
        ps.fused_levels, ps.coarse_chain, ps.wide_classes = fused, chain, wide
        ps.configure()
        z = torch.empty(r.size, dtype=torch.float64, device="cuda")
        s.ctx.mg_apply(torch.as_tensor(r).cuda(), z)
        err = np.abs(z.cpu().numpy() - ref).max() / np.abs(ref).max()
        print(f"{name} sweeps={sweeps} {mode} {precision}: {err:.2e} plan={s.ctx.mg_plan()}")
        assert err <= tol, precision


@pytest.mark.parametrize("name", NAMES)
def test_stretched_trajectory_matches_reference(gpu_available, name):
    """Eight steps of Simulation.run on the GPU vs the reference's (direct
    solver, which the GPU runs as multigrid-CG to 1e-13): velocity within
    1e-10 relative, pressure 1e-8, dt 1e-12, kinetic energy 1e-10."""
    g = golden("stretched")
    tag = f"traj_{name}"
    case = oracle_case(name)
    u0, v0 = initial(name, case)
    sim = _sim(name, u0, v0, pressure_solver="direct", poisson_tol=1e-12)
    res = sim.run(max_steps=8)
    scale = max(np.abs(g[f"{tag}_u"]).max(), 1.0)
    assert np.abs(sim.fields.u - g[f"{tag}_u"]).max() <= 1e-10 * scale
    assert np.abs(sim.fields.v - g[f"{tag}_v"]).max() <= 1e-10 * scale
    pscale = max(np.abs(g[f"{tag}_p"]).max(), 1.0)
    assert np.abs(sim.fields.p - g[f"{tag}_p"]).max() <= 1e-8 * pscale
    np.testing.assert_allclose(res.history["dt"], g[f"{tag}_dt"], rtol=1e-12)
    np.testing.assert_allclose(res.history["kinetic_energy"], g[f"{tag}_kinetic_energy"],
                               rtol=1e-10)
    if name.startswith("s_cyl"):
        np.testing.assert_allclose(res.history["fx"], g[f"{tag}_fx"], rtol=1e-8, atol=1e-10)
    assert max(res.history["max_div"]) <= max(1e-12, 10 * g[f"{tag}_max_div"].max())


def test_stretched_periodic_channel_matches_oracle(gpu_available):
    """A uniform periodic x axis beside a stretched y axis (allowed by
    solver.py:209-230), body-force channel: 8 steps vs the oracle, velocity
    within 1e-10 relative, divergence <= 1e-12."""
    name = "s_pchannel"
    case = oracle_case(name, pressure_solver="direct", tol=1e-12)
    u0, v0 = initial(name, case)
    u, v, p, hist, _ = O.run(case, 8, u0, v0)
    sim = _sim(name, u0, v0, pressure_solver="direct", poisson_tol=1e-12)
    res = sim.run(max_steps=8)
    scale = max(np.abs(u).max(), 1.0)
    assert np.abs(sim.fields.u - u).max() <= 1e-10 * scale
    assert np.abs(sim.fields.v - v).max() <= 1e-10 * scale
    np.testing.assert_allclose(res.history["dt"], hist["dt"], rtol=1e-12)
    np.testing.assert_allclose(res.history["kinetic_energy"], hist["kinetic_energy"],
                               rtol=1e-10)
    assert max(res.history["max_div"]) <= 1e-12


@pytest.mark.parametrize("ps", ["cg", "sor", "multigrid", "mgcg", "jacobi"])
def test_stretched_every_solver_gives_the_same_step(gpu_available, ps):
    """test_solver.py:106-119 on a stretched cavity: every pressure solver at
    a tight tolerance agrees with the direct one to 1e-9 after three steps."""
    kw = (dict(poisson_tol=1e-11, poisson_maxiter=400000) if ps == "jacobi"
          else dict(poisson_tol=1e-13, poisson_maxiter=200000))
    a = _sim("s_cavity", pressure_solver="direct")
    b = _sim("s_cavity", pressure_solver=ps, **kw)
    a.run(max_steps=3)
    b.run(max_steps=3)
    assert np.abs(a.fields.u - b.fields.u).max() < 1e-9
    assert np.abs(a.fields.v - b.fields.v).max() < 1e-9


@pytest.mark.parametrize("name", ["s_cavity", "s_cyl"])
def test_stretched_les_viscosity_and_rhs_bitwise(gpu_available, name):
    """Smagorinsky on a stretched mesh: per-cell filter width sqrt(hx hy),
    array-metric strain rates (turbulence.py:40-113) and the stress-form RHS
    (solver.py:467-529), bit for bit against the reference."""
    g = golden("les_stretched")
    s = _solver(name, use_les=True, smagorinsky_cs=0.17)
    nu_c, nu_k = s.turbulence.eddy_viscosity(g[f"{name}_u"], g[f"{name}_v"], s.nu)
    assert np.array_equal(nu_c, g[f"{name}_nu_c"])
    assert np.array_equal(nu_k, g[f"{name}_nu_corner"])
    for b in (0, 7):
        ru, rv = s.momentum_rhs(g[f"{name}_u"], g[f"{name}_v"], b / 10)
        assert np.array_equal(ru, g[f"{name}_b{b}_rhs_u"]), b
        assert np.array_equal(rv, g[f"{name}_b{b}_rhs_v"]), b


@pytest.mark.parametrize("name", ["s_cavity", "s_cyl"])
def test_stretched_les_trajectory_matches_reference(gpu_available, name):
    """Eight LES steps on a stretched mesh vs the reference: velocity within
    1e-10 relative (direct solve at 1e-13), dt 1e-12."""
    g = golden("les_stretched")
    case = oracle_case(name)
    u0, v0 = initial(name, case)
    sim = _sim(name, u0, v0, pressure_solver="direct", poisson_tol=1e-12, use_les=True,
               smagorinsky_cs=0.17)
    res = sim.run(max_steps=8)
    scale = max(np.abs(g[f"traj_{name}_u"]).max(), 1.0)
    assert np.abs(sim.fields.u - g[f"traj_{name}_u"]).max() <= 1e-10 * scale
    assert np.abs(sim.fields.v - g[f"traj_{name}_v"]).max() <= 1e-10 * scale
    np.testing.assert_allclose(res.history["dt"], g[f"traj_{name}_dt"], rtol=1e-12)
    np.testing.assert_allclose(res.history["kinetic_energy"], g[f"traj_{name}_kinetic_energy"],
                               rtol=1e-10)
