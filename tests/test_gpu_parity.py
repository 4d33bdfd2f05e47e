"""Parity of the HIP engine with the reference (golden vectors) and the oracle.

Bit-exact where the arithmetic is fixed (stencils, boundary passes); to the
stated tolerance where iterative solvers or reductions are involved.
Tolerances are written in each test.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases_def import CASES, initial, oracle_case, product_config
from tests.conftest import golden

pytestmark = pytest.mark.gpu


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


# ------------------------------------------------------------ stencils
@pytest.mark.parametrize("name", ["cavity", "periodic"])
@pytest.mark.parametrize("blend", [0, 7])
def test_momentum_rhs_bitwise(gpu_available, name, blend):
    g = golden("momentum_rhs")
    tag = f"{name}_b{blend}"
    s = _solver(name)
    ru, rv = s.momentum_rhs(g[f"{tag}_u"], g[f"{tag}_v"], blend / 10.0)
    assert np.array_equal(ru, g[f"{tag}_rhs_u"])
    assert np.array_equal(rv, g[f"{tag}_rhs_v"])


@pytest.mark.parametrize("name", ["cavity", "periodic"])
def test_divergence_bitwise(gpu_available, name):
    g = golden("momentum_rhs")
    s = _solver(name)
    assert np.array_equal(s.divergence(g[f"{name}_b0_u"], g[f"{name}_b0_v"]), g[f"{name}_div"])


# ------------------------------------------------------------ boundaries
@pytest.mark.parametrize("name", list(CASES))
def test_boundary_passes(gpu_available, name):
    """apply_velocity bit-exact; the predictor pass (mass balance + IB mask +
    wrap) to 1e-14 and its force to 1e-12 (tree vs pairwise sums)."""
    from pycfd_amd.core.fields import to_device, to_host
    g = golden("boundary")
    s = _solver(name)
    ctx = s.ctx
    for pred in (1, 0):
        u, v = to_device(g[f"{name}_u"]), to_device(g[f"{name}_v"])
        ctx.apply_velocity_bc(u, v, bool(pred))
        assert np.array_equal(to_host(u), g[f"{name}_bc{pred}_u"]), pred
        assert np.array_equal(to_host(v), g[f"{name}_bc{pred}_v"]), pred
    u, v = to_device(g[f"{name}_u"]), to_device(g[f"{name}_v"])
    ctx.apply_velocity_bc(u, v, True)
    ctx.mass_balance(u, v)
    ctx.mask_obstacle(u, v, 0.01, 1)
    ctx.wrap_faces(u, v)
    np.testing.assert_allclose(to_host(u), g[f"{name}_pass_u"], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(to_host(v), g[f"{name}_pass_v"], rtol=1e-14, atol=1e-14)
    if s.has_obstacle:
        np.testing.assert_allclose(ctx.force()[:2], g[f"{name}_pass_force"], rtol=1e-12,
                                   atol=1e-12)
    p = to_device(g[f"{name}_p_in"])
    ctx.apply_pressure_bc(p)
    assert np.array_equal(to_host(p), g[f"{name}_p_bc"])


# ------------------------------------------------------------ Poisson
@pytest.mark.parametrize("name", ["cavity", "periodic", "cyl_pout"])
def test_operator_apply(gpu_available, name):
    import torch
    s = _solver(name)
    A = s.poisson.A
    x = np.random.default_rng(5).standard_normal(A.shape[0])
    y = torch.empty(A.shape[0], dtype=torch.float64, device="cuda")
    s.ctx.poisson_apply(torch.as_tensor(x).cuda(), y)
    ref = A @ x
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-14 * np.abs(ref).max()


def _mg_solver(name, fused=True, chain=-1, stream=0, rows=0, **kw):
    """fused: stored levels through the fused kernels where they fit; chain:
    the cooperative launch of the deepest levels (-1 never, k from level k);
    stream: level 0's Krylov-start legs in streaming form (0 measured, -1
    never, k rows per wavefront); rows: pass-by-pass stored levels with the
    register-resident row kernels (0) or the looping ones (-1)."""
    s = _solver(name, pressure_solver=kw.pop("pressure_solver", "mgcg"), **kw)
    ps = s.pressure_solver
    if not fused or chain != 0 or stream != 0 or rows != 0:
        ps.fused_levels = bool(fused)
        ps.coarse_chain = chain
        ps.stream_rows = stream
        ps.row_passes = rows
        ps.configure()
    return s


@pytest.mark.parametrize("name", ["cavity", "periodic", "cyl_outlet", "cyl_pout"])
@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("fused,rows", [(True, 0), (False, 0), (False, -1)],
                         ids=["fused", "passes", "passes_loop"])
def test_vcycle_matches_reference(gpu_available, name, sweeps, fused, rows):
    """z = M r: one V-cycle of the reference hierarchy (coarsest level at 96
    unknowns) with projections, FP64, vs the reference's, 1e-10: fused stored
    levels, pass-by-pass with register-resident rows, and with looping rows."""
    import torch
    g = golden("multigrid")
    tag = f"{name}_s{sweeps}"
    s = _mg_solver(name, fused, -1, rows=rows, mg_sweeps=sweeps, mg_coarse_max=96)
    r = torch.as_tensor(g[f"{tag}_r"]).cuda()
    z = torch.empty_like(r)
    s.ctx.mg_apply(r, z)
    ref = g[f"{tag}_z"]
    assert np.abs(z.cpu().numpy() - ref).max() <= 1e-10 * np.abs(ref).max()


VCYCLE_GRIDS = [("cavity", {}), ("periodic", {}), ("cyl_outlet", {}), ("cyl_pout", {}),
                ("cyl_outlet", dict(nx=256, ny=128)), ("cyl_pout", dict(nx=160, ny=72)),
                ("periodic", dict(nx=128, ny=96)), ("cavity", dict(nx=37, ny=23))]


CYCLE_MODES = {"fused": (True, -1, -1, 0), "passes": (False, -1, -1, 0),
               "passes_loop": (False, -1, -1, -1), "chain1": (True, 1, -1, 0),
               "chain2": (True, 2, -1, 0), "stream8": (True, -1, 8, 0),
               "stream16": (True, -1, 16, 0)}


@pytest.mark.parametrize("name,grid", VCYCLE_GRIDS,
                         ids=[f"{n}-{g.get('nx', 'g')}" for n, g in VCYCLE_GRIDS])
@pytest.mark.parametrize("coarse_max", [96, 2048])
@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("mode", list(CYCLE_MODES))
def test_vcycle_matches_oracle(gpu_available, name, grid, coarse_max, sweeps, mode):
    """One V-cycle with the coarsest solve at the first level of at most
    coarse_max unknowns, vs the oracle's cycle on the same hierarchy:
    FP64 to 1e-10, the FP32 cycle to 1e-4 (relative to max|z|).  Modes: fused
    level kernels, pass-by-pass levels (register-resident or looping row
    kernels), the cooperative chain from level 1 or 2, level 0's streaming
    legs (8 or 16 rows per wavefront; one sweep only)."""
    fused, chain, stream, rows = CYCLE_MODES[mode]
    if stream > 0 and sweeps != 1:
        pytest.skip("the streaming legs serve one sweep each way")
    import torch
    case = oracle_case(name, **grid)
    S = O.assemble(case)
    H = O.Hierarchy(S.A, (case.nx, case.ny), S.fluid, S.singular, sweeps,
                    coarse_max=coarse_max)
    rng = np.random.default_rng(21)
    r = rng.standard_normal(S.A.shape[0])
    ref = H.apply(r)
    for precision, tol in (("fp64", 1e-10), ("mixed", 1e-4)):
        s = _mg_solver(name, fused, chain, stream, rows, mg_sweeps=sweeps,
                       mg_coarse_max=coarse_max, precision=precision, **grid)
        plan = s.ctx.mg_plan()
        if stream > 0 and len(plan) > 1:  # a one-level hierarchy is the dense solve alone
            assert plan[0]["mode"] == "stream"
        z = torch.empty(r.size, dtype=torch.float64, device="cuda")
        s.ctx.mg_apply(torch.as_tensor(r).cuda(), z)
        err = np.abs(z.cpu().numpy() - ref).max() / np.abs(ref).max()
        print(f"{name} {grid} coarse_max={coarse_max} sweeps={sweeps} {mode} "
              f"{precision}: {err:.2e} plan={s.ctx.mg_plan()}")
        assert err <= tol, precision


@pytest.mark.parametrize("name", ["cavity", "periodic", "cyl_outlet", "cyl_pout"])
@pytest.mark.parametrize("kind", ["direct", "cg", "sor", "multigrid", "mgcg", "jacobi"])
def test_pressure_solvers_match_reference(gpu_available, name, kind):
    """Solution of lap(p) = rhs to 1e-7 relative of the reference's output
    (both solved to relative residual 1e-10); direct to 1e-10."""
    g = golden("multigrid")
    tag = f"{name}_s1"
    key = f"{tag}_{kind}_p"
    if key not in g:
        pytest.skip("no reference output recorded for this pair")
    s = _solver(name, pressure_solver=kind, poisson_tol=1e-10, poisson_maxiter=200000)
    p = s.pressure_solver.solve(g[f"{tag}_rhs"])
    ref = g[key]
    tol = 1e-10 if kind == "direct" else 1e-7
    assert np.abs(p - ref).max() <= tol * np.abs(ref).max()
    assert s.pressure_solver.last_residual < (1e-12 if kind == "direct" else 1e-10)


def test_solver_failure_is_raised(gpu_available):
    """Forced from the host: a CG budget of 5 iterations cannot reach 1e-10."""
    s = _solver("cavity", pressure_solver="cg", poisson_maxiter=5)
    rhs = np.random.default_rng(2).standard_normal((24, 20))
    with pytest.raises(RuntimeError, match="failed to converge"):
        s.pressure_solver.solve(rhs)
    s = _solver("cavity", pressure_solver="sor", poisson_maxiter=5)
    with pytest.raises(RuntimeError, match="stalled"):
        s.pressure_solver.solve(rhs)


@pytest.mark.parametrize("kind", ["cg", "mgcg"])
def test_zero_rhs_after_a_solve(gpu_available, kind):
    """A zero right-hand side right after a solve that sized the iteration
    chunk: scipy's cg returns b itself for ||b|| = 0 (pressure.py:389, 620),
    so the pressure is zero after no iterations; every iteration the host
    enqueued before polling must be a no-op."""
    s = _solver("cavity", pressure_solver=kind, poisson_tol=1e-10, poisson_maxiter=2000)
    rhs = np.random.default_rng(5).standard_normal((24, 20))
    rhs -= rhs.mean()
    s.pressure_solver.solve(rhs)
    assert s.pressure_solver.last_iterations > 0
    p = s.pressure_solver.solve(np.zeros_like(rhs))
    assert s.pressure_solver.last_iterations == 0
    assert np.all(p == 0.0)


# ------------------------------------------------------------ whole steps
TRAJ = [("cavity", "direct", 20), ("cavity", "cg", 5), ("periodic", "direct", 10),
        ("channel", "direct", 10), ("cyl_outlet", "direct", 10), ("cyl_pout", "direct", 10)]


@pytest.mark.parametrize("name,ps,steps", TRAJ)
def test_trajectory_matches_reference(gpu_available, name, ps, steps):
    """Simulation.run on the GPU vs the reference: velocity within 1e-10
    relative, pressure within 1e-8, divergence <= 1e-12 after every step."""
    g = golden("trajectory")
    tag = f"{name}_{ps}"
    case = oracle_case(name)
    u0, v0 = initial(name, case)
    sim = _sim(name, u0, v0, pressure_solver=ps, poisson_tol=1e-12)
    np.testing.assert_allclose(sim.fields.u, g[f"{tag}_u0"], rtol=0, atol=1e-11)
    res = sim.run(max_steps=steps)
    scale = max(np.abs(g[f"{tag}_u"]).max(), 1.0)
    assert np.abs(sim.fields.u - g[f"{tag}_u"]).max() <= 1e-10 * scale
    assert np.abs(sim.fields.v - g[f"{tag}_v"]).max() <= 1e-10 * scale
    pscale = max(np.abs(g[f"{tag}_p"]).max(), 1.0)
    assert np.abs(sim.fields.p - g[f"{tag}_p"]).max() <= 1e-8 * pscale
    np.testing.assert_allclose(res.history["dt"], g[f"{tag}_dt"], rtol=1e-12)
    np.testing.assert_allclose(res.history["kinetic_energy"], g[f"{tag}_kinetic_energy"],
                               rtol=1e-10)
    if name.startswith("cyl"):
        np.testing.assert_allclose(res.history["fx"], g[f"{tag}_fx"], rtol=1e-8, atol=1e-10)
    assert max(res.history["max_div"]) <= 1e-12


@pytest.mark.parametrize("ps", ["sor", "multigrid", "mgcg", "jacobi", "multigrid-96",
                                "mgcg-96", "mgcg-40", "multigrid-chain", "mgcg-chain"])
def test_every_solver_gives_the_same_step(gpu_available, ps):
    """test_solver.py:106-119: every pressure solver at tol 1e-13 agrees with
    the direct one to 1e-9 after a few steps (suffix: mg_coarse_max)."""
    ps, _, cm = ps.partition("-")
    if ps == "jacobi":
        kw = dict(poisson_tol=1e-11, poisson_maxiter=400000)
    else:
        kw = dict(poisson_tol=1e-13, poisson_maxiter=200000)
    chain = cm == "chain"
    if cm:
        kw["mg_coarse_max"] = 40 if chain else int(cm)
    a = _sim("cavity", pressure_solver="direct")
    b = _sim("cavity", pressure_solver=ps, **kw)
    if chain:
        b.solver.pressure_solver.coarse_chain = 1
        b.solver.pressure_solver.configure()
        assert any(p["mode"] == "chain" for p in b.solver.ctx.mg_plan())
    a.run(max_steps=3)
    b.run(max_steps=3)
    assert np.abs(a.fields.u - b.fields.u).max() < 1e-9
    assert np.abs(a.fields.v - b.fields.v).max() < 1e-9


def test_obstacle_faces_exactly_zero(gpu_available):
    sim = _sim("cyl_outlet", 1.0, None, pressure_solver="mgcg")
    sim.run(max_steps=3)
    s = sim.solver
    u = sim.fields.u[s.u_upd]
    v = sim.fields.v[s.v_upd]
    assert np.all(u[s.u_face_solid] == 0.0) and np.all(v[s.v_face_solid] == 0.0)


def test_quiescent_stays_quiescent(gpu_available):
    sim = _sim("periodic", pressure_solver="mgcg")
    sim.run(max_steps=3)
    assert np.abs(sim.fields.u).max() == 0.0 and np.abs(sim.fields.v).max() == 0.0


def test_run_twice_bitwise(gpu_available):
    """Deterministic reductions: two identical runs agree bit for bit."""
    a = _sim("cyl_outlet", 1.0, None, pressure_solver="mgcg")
    b = _sim("cyl_outlet", 1.0, None, pressure_solver="mgcg")
    a.run(max_steps=4)
    b.run(max_steps=4)
    assert np.array_equal(a.fields.u, b.fields.u) and np.array_equal(a.fields.p, b.fields.p)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("coarse_max", [96, 2048])
@pytest.mark.parametrize("stream", [-1, 8])
def test_oracle_parity_mgcg(gpu_available, precision, coarse_max, stream):
    """mgcg at tol 1e-12 vs the oracle's direct solve, cylinder with OUTLET,
    with level 0's tiled or streaming legs."""
    case = oracle_case("cyl_outlet", pressure_solver="direct")
    u0, v0 = initial("cyl_outlet", case)
    u, v, p, hist, t = O.run(case, 6, u0, v0)
    sim = _sim("cyl_outlet", u0, v0, pressure_solver="mgcg", poisson_tol=1e-12,
               poisson_maxiter=300, precision=precision, mg_coarse_max=coarse_max)
    ps = sim.solver.pressure_solver
    ps.stream_rows = stream
    ps.configure()
    sim.run(max_steps=6)
    assert np.abs(sim.fields.u - u).max() <= 1e-9 * np.abs(u).max()
    assert np.abs(sim.fields.v - v).max() <= 1e-9 * np.abs(u).max()


# ------------------------------------------------------------ checkpoints
@pytest.mark.parametrize("ps", ["mgcg", "sor"])
def test_resume_from_checkpoint_is_bitwise(gpu_available, tmp_path, ps):
    """Six steps straight through == three steps, checkpoint, resume, three
    steps: bit for bit, because the warm-start pressure travels with the state
    (the reference's resume restarts its iterative solvers cold)."""
    from pycfd_amd.physics.incompressible import Simulation
    kw = dict(pressure_solver=ps, poisson_tol=1e-10, mg_coarse_max=96)
    case = oracle_case("cyl_outlet")
    u0, v0 = initial("cyl_outlet", case)
    a = _sim("cyl_outlet", u0, v0, **kw)
    a.run(max_steps=6)
    b = _sim("cyl_outlet", u0, v0, **kw)
    b.run(max_steps=3)
    path = b.save_checkpoint(tmp_path / "state")
    c = Simulation.from_checkpoint(path, obstacle=b.obstacle)
    assert c.fields.step == 3 and c.fields.t == b.fields.t
    c.run(max_steps=3)
    assert c.fields.step == 6
    assert np.array_equal(c.fields.u, a.fields.u)
    assert np.array_equal(c.fields.v, a.fields.v)
    assert np.array_equal(c.fields.p, a.fields.p)


# ------------------------------------------------------------ Smagorinsky LES
@pytest.mark.parametrize("name", ["cavity", "periodic"])
def test_les_viscosity_and_rhs_bitwise(gpu_available, name):
    """Eddy viscosity (turbulence.py:63-113) and the stress-form momentum RHS
    (solver.py:467-529) on random fields: bit for bit against the reference."""
    g = golden("les")
    s = _solver(name, use_les=True, smagorinsky_cs=0.17)
    nu_c, nu_k = s.turbulence.eddy_viscosity(g[f"{name}_u"], g[f"{name}_v"], s.nu)
    assert np.array_equal(nu_c, g[f"{name}_nu_c"])
    assert np.array_equal(nu_k, g[f"{name}_nu_corner"])
    for b in (0, 7):
        ru, rv = s.momentum_rhs(g[f"{name}_u"], g[f"{name}_v"], b / 10)
        assert np.array_equal(ru, g[f"{name}_b{b}_rhs_u"])
        assert np.array_equal(rv, g[f"{name}_b{b}_rhs_v"])


def test_les_closed_forms(gpu_available):
    """test_solver.py:262-289: a uniform periodic flow has nu_c = nu; with a
    constant viscosity the stress form equals nu * lap on a solenoidal field."""
    s = _solver("periodic", use_les=True)
    f = s.initialize(1.0, -0.5)
    nu_c, _ = s.turbulence.eddy_viscosity(f.u, f.v, s.nu)
    assert np.abs(nu_c - s.nu).max() < 1e-14
    lam = _solver("periodic")
    Xu, Yu = lam.mesh.u_grid()
    Xv, Yv = lam.mesh.v_grid()
    f = lam.initialize(-np.cos(Xu) * np.sin(Yu), np.sin(Xv) * np.cos(Yv))
    nx, ny = lam.mesh.shape
    ru, rv = lam.momentum_rhs(f.u, f.v)
    rvu, rvv = lam.momentum_rhs_variable(f.u, f.v, np.full((nx + 2, ny + 2), lam.nu))
    assert np.abs(ru - rvu).max() < 1e-14
    assert np.abs(rv - rvv).max() < 1e-14


@pytest.mark.parametrize("name,steps", [("cavity", 10), ("cyl_outlet", 6)])
def test_les_trajectory_matches_reference(gpu_available, name, steps):
    """Whole LES steps through Simulation.run vs the reference, velocity to
    1e-10 relative (direct solve at 1e-13)."""
    g = golden("les")
    case = oracle_case(name)
    u0, v0 = initial(name, case)
    sim = _sim(name, u0, v0, pressure_solver="direct", use_les=True, smagorinsky_cs=0.17)
    sim.run(max_steps=steps)
    scale = max(np.abs(g[f"traj_{name}_u"]).max(), 1.0)
    assert np.abs(sim.fields.u - g[f"traj_{name}_u"]).max() <= 1e-10 * scale
    assert np.abs(sim.fields.v - g[f"traj_{name}_v"]).max() <= 1e-10 * scale
    assert max(sim.result.history["max_div"]) <= 1e-11


def test_tune_file_records_measured_plan(gpu_available, tmp_path, monkeypatch):
    """With PCFD_TUNE_FILE set, each measured level plan is appended to the
    file (14 key and 4 choice integers a line), and a solver configured again
    for the same operator takes the same plan."""
    f = tmp_path / "tune.txt"
    monkeypatch.setenv("PCFD_TUNE_FILE", str(f))
    kw = dict(nx=44, ny=36, mg_coarse_max=96)  # a shape no other test tunes
    s = _mg_solver("cavity", **kw)
    rows = [list(map(int, line.split())) for line in f.read_text().splitlines() if line.strip()]
    assert rows and all(len(r) == 18 for r in rows)
    assert any(r[0] == 1 and r[1:3] == [44, 36] for r in rows), rows  # level 0 of this grid
    plan = s.ctx.mg_plan()
    assert len(plan) > 1
    assert _mg_solver("cavity", **kw).ctx.mg_plan() == plan


# ------------------------------------------------------------ fused passes, power-of-two spacing
@pytest.mark.parametrize("name", ["cavity", "periodic", "channel", "cyl_outlet", "cyl_pout"])
def test_projection_ghost_ring_equals_apply_pressure(gpu_available, name):
    """The correction kernel fills the pressure ghost layer itself: the ring
    it writes equals the oracle's apply_pressure (boundary.py:166-169,
    293-303, 412-420, pinned bit for bit to the reference) applied to the
    same interior, corners included, and the velocity correction equals the
    oracle's on the same pressure."""
    from pycfd_amd.core.fields import device_zeros, to_device, to_host
    rng = np.random.default_rng(11)
    s = _solver(name, pressure_solver="mgcg", poisson_tol=1e-12)
    case = oracle_case(name)
    nx, ny = case.nx, case.ny
    u0 = rng.standard_normal((nx + 3, ny + 2))
    v0 = rng.standard_normal((nx + 2, ny + 3))
    ud, vd = to_device(u0), to_device(v0)
    p = device_zeros((nx + 2, ny + 2))
    dt = 0.0123
    s.ctx.project(ud, vd, dt, p)
    ph = to_host(p)
    want = np.zeros_like(ph)
    want[1:nx + 1, 1:ny + 1] = ph[1:nx + 1, 1:ny + 1]
    O.apply_pressure(case, want)
    assert np.array_equal(ph, want)
    u_lo, u_hi, v_lo, v_hi = O.upd_ranges(case)
    uw, vw = u0.copy(), v0.copy()
    uw[u_lo:u_hi, 1:ny + 1] -= dt * (want[u_lo:u_hi, 1:ny + 1] - want[u_lo - 1:u_hi - 1, 1:ny + 1]) / case.dx
    vw[1:nx + 1, v_lo:v_hi] -= dt * (want[1:nx + 1, v_lo:v_hi] - want[1:nx + 1, v_lo - 1:v_hi - 1]) / case.dy
    assert np.array_equal(to_host(ud), uw) and np.array_equal(to_host(vd), vw)


POW2 = [("cavity", dict(nx=32, ny=16, lx=1.0, ly=0.5)),
        ("periodic", dict(nx=32, ny=16, lx=2.0, ly=1.0)),
        ("cyl_outlet", dict(nx=64, ny=32, lx=16.0, ly=8.0))]


@pytest.mark.parametrize("name,grid", POW2, ids=[p[0] for p in POW2])
@pytest.mark.parametrize("blend", [0.0, 0.7])
def test_power_of_two_spacing_bitwise(gpu_available, name, grid, blend):
    """On a grid whose spacings are powers of two the kernels multiply by the
    exact reciprocals instead of dividing: the momentum RHS and divergence
    stay bit-identical to the oracle's NumPy arithmetic (itself pinned bit for
    bit to the reference), and so does one Euler predictor step."""
    from pycfd_amd.core.fields import to_device, to_host
    rng = np.random.default_rng(5)
    s = _solver(name, **grid)
    assert s.ctx is not None
    case = oracle_case(name, **grid)
    assert case.dx in (2.0 ** -5, 2.0 ** -4, 2.0 ** -2) and case.dy == case.dx
    u = rng.standard_normal(s.mesh.u_shape)
    v = rng.standard_normal(s.mesh.v_shape)
    ru, rv = s.momentum_rhs(u, v, blend)
    wu, wv = O.momentum_rhs(case, u, v, blend)
    assert np.array_equal(ru, wu) and np.array_equal(rv, wv)
    assert np.array_equal(s.divergence(u, v), O.divergence(case, u, v))


@pytest.mark.parametrize("name,grid", POW2, ids=[p[0] for p in POW2])
def test_power_of_two_trajectory_matches_oracle(gpu_available, name, grid):
    """Whole RK3 steps on a power-of-two grid against the oracle: velocity
    within 1e-10 relative, pressure within 1e-8."""
    case = oracle_case(name, pressure_solver="direct", tol=1e-12, **grid)
    u0, v0 = initial(name, case)
    sim = _sim(name, u0, v0, pressure_solver="mgcg", poisson_tol=1e-13, **grid)
    sim.run(max_steps=6)
    u, v, p, hist, t = O.run(case, 6, u0, v0)
    scale = max(np.abs(u).max(), 1.0)
    assert np.abs(sim.fields.u - u).max() <= 1e-10 * scale
    assert np.abs(sim.fields.v - v).max() <= 1e-10 * scale
    assert np.abs(sim.fields.p - p).max() <= 1e-8 * max(np.abs(p).max(), 1.0)
