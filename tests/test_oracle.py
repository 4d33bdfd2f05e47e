"""Pin the CPU oracle before trusting it as the checker of the HIP engine.

Each oracle function is checked against the golden vectors recorded from the
reference (tests/golden/make_golden.py) and, where the mathematics fixes the
answer, against closed forms, dense linear algebra and invariants chosen so
that a dropped term, a sign or index slip, or a transposed operand fails.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from tests.cases_def import CASES, circle, initial, oracle_case
from tests.conftest import golden


# ---------------------------------------------------------------- stencils
@pytest.mark.parametrize("name", ["cavity", "periodic"])
@pytest.mark.parametrize("blend", [0, 7])
def test_momentum_rhs_matches_reference_bitwise(name, blend):
    g = golden("momentum_rhs")
    tag = f"{name}_b{blend}"
    case = oracle_case(name)
    ru, rv = O.momentum_rhs(case, g[f"{tag}_u"], g[f"{tag}_v"], blend / 10.0)
    assert np.array_equal(ru, g[f"{tag}_rhs_u"])
    assert np.array_equal(rv, g[f"{tag}_rhs_v"])


@pytest.mark.parametrize("name", ["cavity", "periodic"])
def test_divergence_matches_reference_bitwise(name):
    g = golden("momentum_rhs")
    case = oracle_case(name)
    assert np.array_equal(O.divergence(case, g[f"{name}_b0_u"], g[f"{name}_b0_v"]),
                          g[f"{name}_div"])


def test_momentum_rhs_closed_forms():
    """Uniform flow: no advection, no diffusion -> rhs = body force exactly;
    a linear shear u = y is steady (lap = 0, d(uu)/dx = 0, d(uv)/dy = 0 with v = 0)."""
    case = O.Case(nx=8, ny=6, lx=1.0, ly=1.0, body_force=(0.25, -0.5))
    u = np.full((11, 8), 3.0)
    v = np.full((10, 9), -2.0)
    ru, rv = O.momentum_rhs(case, u, v)
    assert np.all(ru == 0.25) and np.all(rv == -0.5)
    case = O.Case(nx=8, ny=6, lx=1.0, ly=1.0)
    y = (np.arange(8) - 0.5) / 6.0
    u = np.tile(y, (11, 1))
    ru, rv = O.momentum_rhs(case, u, np.zeros((10, 9)))
    assert np.abs(ru).max() < 1e-12 and np.abs(rv).max() < 1e-12


# ---------------------------------------------------------------- boundaries
@pytest.mark.parametrize("name", list(CASES))
def test_boundary_passes_match_reference(name):
    g = golden("boundary")
    case = oracle_case(name)
    for pred in (1, 0):
        u, v = g[f"{name}_u"].copy(), g[f"{name}_v"].copy()
        O.apply_velocity(case, u, v, bool(pred))
        assert np.array_equal(u, g[f"{name}_bc{pred}_u"])
        assert np.array_equal(v, g[f"{name}_bc{pred}_v"])
    u, v = g[f"{name}_u"].copy(), g[f"{name}_v"].copy()
    st = O.State()
    O.apply_velocity(case, u, v, True)
    O.mass_balance(case, u, v)
    O.mask(case, st, u, v, 0.01, reset=True)
    O.wrap_faces(case, u, v)
    np.testing.assert_allclose(u, g[f"{name}_pass_u"], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(v, g[f"{name}_pass_v"], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(st.force, g[f"{name}_pass_force"], rtol=1e-12, atol=1e-12)
    p = g[f"{name}_p_in"].copy()
    O.apply_pressure(case, p)
    assert np.array_equal(p, g[f"{name}_p_bc"])


def test_mass_balance_closes_the_flux():
    """After the balance the net boundary flux of a random field vanishes."""
    case = oracle_case("channel")
    rng = np.random.default_rng(0)
    u = rng.random((case.nx + 3, case.ny + 2))
    v = rng.random((case.nx + 2, case.ny + 3))
    O.apply_velocity(case, u, v, True)
    O.mass_balance(case, u, v)
    nx, ny = case.nx, case.ny
    net = (u[nx + 1, 1:ny + 1].sum() - u[1, 1:ny + 1].sum()) * case.dy + \
          (v[1:nx + 1, ny + 1].sum() - v[1:nx + 1, 1].sum()) * case.dx
    assert abs(net) < 1e-12


def test_circle_mask_matches_reference():
    g = golden("circle_mask")
    for nx, ny, lx, ly, cx in ((2048, 1024, 32.0, 16.0, 4.0), (512, 256, 16.0, 8.0, 4.0),
                                (48, 24, 12.0, 6.0, 3.0)):
        m = circle(nx, ny, lx, ly, cx, ly / 2, 0.5)
        bits = np.unpackbits(g[f"m{nx}x{ny}"])[: nx * ny].astype(bool).reshape(nx, ny)
        assert np.array_equal(m, bits)


# ---------------------------------------------------------------- operator
@pytest.mark.parametrize("name", list(CASES))
def test_operator_invariants(name):
    case = oracle_case(name)
    S = O.assemble(case)
    A = S.A
    assert abs(A - A.T).max() == 0.0                       # symmetric
    if S.singular:                                         # null space = constant
        rs = np.asarray(A.sum(axis=1)).ravel()[S.fluid]
        assert np.abs(rs).max() < 1e-9 * abs(A.diagonal()).max()
    assert (A.getnnz(axis=1) <= 5).all()


def test_operator_5point_coefficients():
    case = O.Case(nx=6, ny=5, lx=3.0, ly=1.0)
    A = O.assemble(case).A.toarray()
    cx, cy = 1.0 / (case.dx * case.dx), 1.0 / (case.dy * case.dy)
    k = 2 * 5 + 2                                          # interior cell (2, 2)
    assert A[k, k] == pytest.approx(-(2 * cx + 2 * cy), rel=1e-15)
    assert A[k, k - 5] == cx and A[k, k + 5] == cx and A[k, k - 1] == cy and A[k, k + 1] == cy
    k = 0                                                  # corner: two faces
    assert A[k, k] == pytest.approx(-(cx + cy), rel=1e-15)


@pytest.mark.parametrize("kind", ["direct", "cg", "sor", "jacobi", "multigrid", "mgcg"])
def test_solvers_manufactured_solution(kind):
    """lap(p) = -2 pi^2 cos(pi x) cos(pi y) on the unit square, Neumann:
    the discrete solution converges to cos(pi x)cos(pi y) (test_pressure.py:95-121)."""
    errs = []
    for n in (16, 32):
        case = O.Case(nx=n, ny=n, pressure_solver=kind, tol=1e-11, maxiter=200000)
        X, Y = np.meshgrid(case.xc, case.yc, indexing="ij")
        exact = np.cos(np.pi * X) * np.cos(np.pi * Y)
        p = O.solve(case, O.State(), -2 * np.pi ** 2 * exact)
        errs.append(np.abs(p - (exact - exact.mean())).max())
    assert errs[0] < 5e-3
    assert np.log2(errs[0] / errs[1]) > 1.8


@pytest.mark.parametrize("name", ["cavity", "periodic", "cyl_outlet", "cyl_pout"])
@pytest.mark.parametrize("sweeps", [1, 2])
def test_solvers_match_reference_outputs(name, sweeps):
    g = golden("multigrid")
    tag = f"{name}_s{sweeps}"
    for kind in ("direct", "cg", "sor", "multigrid", "mgcg", "jacobi"):
        key = f"{tag}_{kind}_p"
        if key not in g:
            continue
        case = oracle_case(name, pressure_solver=kind, tol=1e-10, maxiter=200000,
                           mg_sweeps=sweeps)
        st = O.State()
        p = O.solve(case, st, g[f"{tag}_rhs"])
        scale = np.abs(g[key]).max()
        tol = 1e-12 if kind == "direct" else 1e-7
        assert np.abs(p - g[key]).max() <= tol * scale, kind
        if kind in ("direct", "mgcg", "multigrid", "sor"):
            assert st.last_iterations == int(g[f"{tag}_{kind}_it"]), kind


@pytest.mark.parametrize("name", ["cavity", "periodic", "cyl_outlet", "cyl_pout"])
def test_hierarchy_matches_reference(name):
    g = golden("multigrid")
    for sweeps in (1, 2):
        tag = f"{name}_s{sweeps}"
        case = oracle_case(name)
        S = O.assemble(case)
        H = O.Hierarchy(S.A, (case.nx, case.ny), S.fluid, S.singular, sweeps)
        assert [lv["A"].shape[0] for lv in H.levels] == list(g[f"{tag}_unknowns"])
        assert [lv["A"].nnz for lv in H.levels] == list(g[f"{tag}_nnz"])
        np.testing.assert_allclose([lv["weight"] for lv in H.levels], g[f"{tag}_weights"],
                                   rtol=1e-13)
        z = H.apply(g[f"{tag}_r"])
        np.testing.assert_allclose(z, g[f"{tag}_z"], rtol=1e-10,
                                   atol=1e-11 * np.abs(g[f"{tag}_z"]).max())


def test_vcycle_is_symmetric_dense():
    """Materialise M column by column: symmetric to round-off (test_multigrid.py:339-359)."""
    case = oracle_case("cyl_outlet", nx=24, ny=12, lx=6.0, ly=3.0)
    case.solid = circle(24, 12, 6.0, 3.0, 3.0, 1.5, 0.5)
    S = O.assemble(case)
    H = O.Hierarchy(S.A, (24, 12), S.fluid, S.singular, 1)
    n = S.A.shape[0]
    M = np.column_stack([H.apply(np.eye(n)[:, k]) for k in range(n)])
    assert np.abs(M - M.T).max() < 1e-10 * np.abs(M).max()


def test_coarsest_is_pseudo_inverse():
    """A tiny singular Neumann operator: one level, coarse solve = pseudo-inverse."""
    case = O.Case(nx=8, ny=8)
    S = O.assemble(case)
    H = O.Hierarchy(S.A, (8, 8), S.fluid, S.singular, 1)
    assert len(H.levels) == 1
    b = np.random.default_rng(3).standard_normal(64)
    b -= b.mean()
    x = H.apply(b)
    np.testing.assert_allclose(x, np.linalg.pinv(S.A.toarray()) @ b, atol=1e-12)


@pytest.mark.parametrize("name", ["cavity", "periodic", "cyl_pout"])
def test_truncated_hierarchy_whole_level_is_pseudo_inverse(name):
    """coarse_max >= n: the cycle is the exact solve, A z = b on the active
    set with z mean-free (singular) -- checked against a dense lstsq."""
    case = oracle_case(name)
    S = O.assemble(case)
    n = S.A.shape[0]
    H = O.Hierarchy(S.A, (case.nx, case.ny), S.fluid, S.singular, 1, coarse_max=n)
    assert len(H.levels) == 1
    b = np.random.default_rng(4).standard_normal(n)
    z = H.apply(b)
    act = S.fluid
    bp = b.copy()
    bp[~act] = 0.0
    if S.singular:
        bp[act] -= bp[act].mean()
    Aa = S.A.toarray()[np.ix_(act, act)]
    ref = np.zeros(n)
    ref[act] = np.linalg.lstsq(Aa, bp[act], rcond=None)[0]
    if S.singular:
        ref[act] -= ref[act].mean()
    np.testing.assert_allclose(z, ref, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("coarse_max", [96, 200, 400])
def test_truncated_hierarchy_two_grid_properties(coarse_max):
    """Moving the exact solve up a level keeps the preconditioner symmetric,
    and an exact coarse solve contracts at least as well as recursing."""
    case = oracle_case("cyl_outlet", nx=32, ny=24, lx=8.0, ly=6.0)
    case.solid = circle(32, 24, 8.0, 6.0, 3.0, 3.0, 0.5)
    S = O.assemble(case)
    n = S.A.shape[0]
    full = O.Hierarchy(S.A, (32, 24), S.fluid, S.singular, 1)
    H = O.Hierarchy(S.A, (32, 24), S.fluid, S.singular, 1, coarse_max=coarse_max)
    assert [lv["A"].shape[0] for lv in H.levels] == \
        [lv["A"].shape[0] for lv in full.levels][:len(H.levels)]
    assert H.levels[-1]["A"].shape[0] <= max(coarse_max, full.levels[-1]["A"].shape[0])
    I = np.eye(n)
    M = np.column_stack([H.apply(I[:, k]) for k in range(n)])
    assert np.abs(M - M.T).max() < 1e-10 * np.abs(M).max()
    A = S.A.toarray()

    def factor(Mx):
        act = S.fluid
        E = (I - Mx @ A)[np.ix_(act, act)]
        if S.singular:
            Pm = np.eye(act.sum()) - 1.0 / act.sum()
            E = Pm @ E @ Pm
        return np.abs(np.linalg.eigvals(E)).max()

    Mf = np.column_stack([full.apply(I[:, k]) for k in range(n)])
    assert factor(M) <= factor(Mf) + 1e-12
    assert factor(M) < 0.5


# ---------------------------------------------------------------- time steps
TRAJ = [("cavity", "direct", 20), ("cavity", "cg", 5), ("periodic", "direct", 10),
        ("channel", "direct", 10), ("cyl_outlet", "direct", 10), ("cyl_pout", "direct", 10)]


@pytest.mark.parametrize("name,ps,steps", TRAJ)
def test_trajectory_matches_reference(name, ps, steps):
    g = golden("trajectory")
    tag = f"{name}_{ps}"
    case = oracle_case(name, pressure_solver=ps, tol=1e-12)
    u0, v0 = initial(name, case)
    u, v, p, hist, t = O.run(case, steps, u0, v0)
    scale = max(np.abs(g[f"{tag}_u"]).max(), 1.0)
    assert np.abs(u - g[f"{tag}_u"]).max() <= 1e-10 * scale
    assert np.abs(v - g[f"{tag}_v"]).max() <= 1e-10 * scale
    pscale = max(np.abs(g[f"{tag}_p"]).max(), 1.0)
    assert np.abs(p - g[f"{tag}_p"]).max() <= 1e-8 * pscale
    np.testing.assert_allclose(hist["dt"], g[f"{tag}_dt"], rtol=1e-12)
    np.testing.assert_allclose(hist["kinetic_energy"], g[f"{tag}_kinetic_energy"], rtol=1e-10)
    np.testing.assert_allclose(hist["fx"], g[f"{tag}_fx"], rtol=1e-8, atol=1e-10)
    assert max(hist["max_div"]) < 1e-9


def test_taylor_green_decay_rate():
    """Exact solution of the full equations: amplitude decays as exp(-2 nu t)."""
    n, nu, dt, steps = 32, 0.05, 0.01, 10
    case = O.Case(nx=n, ny=n, lx=2 * np.pi, ly=2 * np.pi, re=1 / nu, dt=dt, adaptive_dt=False,
                  kinds=("periodic",) * 4, pressure_solver="direct")
    u0, v0 = initial("periodic", case)
    u, v, p, hist, t = O.run(case, steps, u0, v0)
    Xu, Yu = np.meshgrid(case.xf, case.yc, indexing="ij")
    exact = np.sin(Xu) * np.cos(Yu) * np.exp(-2 * nu * t)
    err = np.abs(u[1:n + 2, 1:n + 1] - exact).max()
    assert err < 5e-3                          # O(h^2) spatial error at 32^2
    ke0 = 0.5 * (2 * np.pi) ** 2 * 0.5
    assert abs(hist["kinetic_energy"][-1] / ke0 / np.exp(-4 * nu * t) - 1) < 2e-2


# ---------------------------------------------------------------- LES closure
@pytest.mark.parametrize("name", ["cavity", "periodic"])
def test_les_viscosity_and_rhs_match_reference_bitwise(name):
    """turbulence.py:63-113 and the stress-form RHS (solver.py:467-529) on
    random fields, against vectors recorded from the reference."""
    g = golden("les")
    case = oracle_case(name, use_les=True)
    nu_c, nu_k = O.eddy_viscosity(case, g[f"{name}_u"], g[f"{name}_v"])
    assert np.array_equal(nu_c, g[f"{name}_nu_c"])
    assert np.array_equal(nu_k, g[f"{name}_nu_corner"])
    for b in (0, 7):
        ru, rv = O.momentum_rhs(case, g[f"{name}_u"], g[f"{name}_v"], b / 10)
        assert np.array_equal(ru, g[f"{name}_b{b}_rhs_u"])
        assert np.array_equal(rv, g[f"{name}_b{b}_rhs_v"])


def test_les_closed_forms():
    """test_solver.py:257-289: a uniform periodic flow has zero strain (nu_c =
    nu); with constant viscosity arrays the stress form is nu * lap on a
    solenoidal field."""
    case = oracle_case("periodic", use_les=True)
    nx, ny = case.nx, case.ny
    st = O.State()
    u, v, _ = O.initialize(case, st, 1.0, -0.5)
    nu_c, _ = O.eddy_viscosity(case, u, v)
    assert np.abs(nu_c - case.nu).max() < 1e-14
    lam = oracle_case("periodic")
    xs = (np.arange(nx + 1) * case.dx)[:, None]
    yc = ((np.arange(ny) + 0.5) * case.dy)[None, :]
    xc = ((np.arange(nx) + 0.5) * case.dx)[:, None]
    ys = (np.arange(ny + 1) * case.dy)[None, :]
    u, v, _ = O.initialize(lam, O.State(), -np.cos(xs) * np.sin(yc), np.sin(xc) * np.cos(ys))
    nuc = np.full((nx + 2, ny + 2), lam.nu)
    nuk = np.full((nx + 1, ny + 1), lam.nu)
    d_u, d_v = O.diffusion_variable(lam, u, v, nuc, nuk)
    ru, rv = O.momentum_rhs(lam, u, v)
    ru_adv, rv_adv = O.momentum_rhs(oracle_case("periodic", re=1e300), u, v)
    np.testing.assert_allclose(d_u, ru - ru_adv, atol=1e-12)
    np.testing.assert_allclose(d_v, rv - rv_adv, atol=1e-12)


@pytest.mark.parametrize("name,steps", [("cavity", 10), ("cyl_outlet", 6)])
def test_les_trajectory_matches_reference(name, steps):
    g = golden("les")
    case = oracle_case(name, pressure_solver="direct", use_les=True)
    u0, v0 = initial(name, case)
    u, v, p, hist, t = O.run(case, steps, u0, v0)
    scale = max(np.abs(g[f"traj_{name}_u"]).max(), 1.0)
    assert np.abs(u - g[f"traj_{name}_u"]).max() <= 1e-10 * scale
    assert np.abs(v - g[f"traj_{name}_v"]).max() <= 1e-10 * scale


# ---------------------------------------------------------------- stretched meshes
STRETCHED_NAMES = ["s_cavity", "s_channel", "s_cyl", "s_cylp"]


@pytest.mark.parametrize("name", STRETCHED_NAMES)
def test_stretched_operators_match_reference(name):
    """mesh.py:122-310 widths and metrics; solver.py:334-465 momentum RHS and
    532-543 divergence bit for bit; the predictor boundary pass and its force;
    pressure.py:193-207 operator, row scale and Dirichlet terms."""
    g = golden("stretched")
    case = oracle_case(name)
    assert np.array_equal(case.dx_cells, g[f"{name}_dx_cells"])
    assert np.array_equal(case.dy_cells, g[f"{name}_dy_cells"])
    u, v = g[f"{name}_u"], g[f"{name}_v"]
    for b in (0, 7):
        ru, rv = O.momentum_rhs(case, u, v, b / 10)
        assert np.array_equal(ru, g[f"{name}_b{b}_rhs_u"]), b
        assert np.array_equal(rv, g[f"{name}_b{b}_rhs_v"]), b
    assert np.array_equal(O.divergence(case, u, v), g[f"{name}_div"])
    us, vs = u.copy(), v.copy()
    st = O.State()
    O.apply_velocity(case, us, vs, True)
    O.mass_balance(case, us, vs)
    O.mask(case, st, us, vs, 0.01, reset=True)
    np.testing.assert_allclose(us, g[f"{name}_pass_u"], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(vs, g[f"{name}_pass_v"], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(st.force, g[f"{name}_pass_force"], rtol=1e-12, atol=1e-14)
    S = O.assemble(case)
    x = g[f"{name}_Ax_x"]
    np.testing.assert_allclose(S.A @ x, g[f"{name}_Ax"], rtol=1e-14, atol=1e-14 *
                               np.abs(g[f"{name}_Ax"]).max())
    np.testing.assert_array_equal(np.broadcast_to(S.row_scale, (x.size,)),
                                  np.broadcast_to(g[f"{name}_row_scale"], (x.size,)))
    np.testing.assert_array_equal(S.bc_rhs, g[f"{name}_bc_rhs"])


def test_stretched_poisson_operator_is_symmetric_and_consistent():
    """pressure.py:137-161: the area-scaled rows make A symmetric; A 1 = 0 on a
    closed stretched box; the scaled rows are div(grad) times area / mean area."""
    case = oracle_case("s_cavity")
    S = O.assemble(case)
    A = S.A.toarray()
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    assert np.abs(A @ np.ones(A.shape[0])).max() < 1e-10 * np.abs(A).max()


@pytest.mark.parametrize("name", STRETCHED_NAMES)
def test_stretched_trajectory_matches_reference(name):
    g = golden("stretched")
    case = oracle_case(name, pressure_solver="direct", tol=1e-12)
    u0, v0 = initial(name, case)
    u, v, p, hist, t = O.run(case, 8, u0, v0)
    scale = max(np.abs(g[f"traj_{name}_u"]).max(), 1.0)
    assert np.abs(u - g[f"traj_{name}_u"]).max() <= 1e-10 * scale
    assert np.abs(v - g[f"traj_{name}_v"]).max() <= 1e-10 * scale
    np.testing.assert_allclose(hist["dt"], g[f"traj_{name}_dt"], rtol=1e-12)
    np.testing.assert_allclose(hist["kinetic_energy"], g[f"traj_{name}_kinetic_energy"],
                               rtol=1e-10)
    np.testing.assert_allclose(hist["fx"], g[f"traj_{name}_fx"], rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("name", ["s_cavity", "s_cyl"])
def test_stretched_les_matches_reference(name):
    """turbulence.py:40-113 with the per-cell filter width and array-metric
    strain rates, solver.py:467-529 stress form: bit for bit; 8 LES steps
    within 1e-10."""
    g = golden("les_stretched")
    case = oracle_case(name, use_les=True, smagorinsky_cs=0.17)
    u, v = g[f"{name}_u"], g[f"{name}_v"]
    nu_c, nu_k = O.eddy_viscosity(case, u, v)
    assert np.array_equal(nu_c, g[f"{name}_nu_c"])
    assert np.array_equal(nu_k, g[f"{name}_nu_corner"])
    for b in (0, 7):
        ru, rv = O.momentum_rhs(case, u, v, b / 10)
        assert np.array_equal(ru, g[f"{name}_b{b}_rhs_u"]), b
        assert np.array_equal(rv, g[f"{name}_b{b}_rhs_v"]), b
    case = oracle_case(name, use_les=True, smagorinsky_cs=0.17, pressure_solver="direct",
                       tol=1e-12)
    u0, v0 = initial(name, case)
    uu, vv, _, hist, _ = O.run(case, 8, u0, v0)
    scale = max(np.abs(g[f"traj_{name}_u"]).max(), 1.0)
    assert np.abs(uu - g[f"traj_{name}_u"]).max() <= 1e-10 * scale
    assert np.abs(vv - g[f"traj_{name}_v"]).max() <= 1e-10 * scale
    np.testing.assert_allclose(hist["dt"], g[f"traj_{name}_dt"], rtol=1e-12)


def test_stretched_les_reduces_to_uniform():
    """A stretched mesh of ratio 1 is the uniform mesh: the array-metric LES
    path evaluated with all-equal widths agrees with the scalar one to
    rounding (a dropped or misplaced metric would not)."""
    case = oracle_case("cavity", use_les=True)
    rng = np.random.default_rng(9)
    u = rng.standard_normal((case.nx + 3, case.ny + 2))
    v = rng.standard_normal((case.nx + 2, case.ny + 3))
    ref_u, ref_v = O.momentum_rhs(case, u, v)
    forced = oracle_case("cavity", use_les=True)
    forced.uniform = False
    nx, ny = forced.nx, forced.ny
    forced.hx, forced.hy = forced.dx_cells[:, None], forced.dy_cells[None, :]
    forced.hxu = np.full((nx + 1, 1), forced.dx)
    forced.hyv = np.full((1, ny + 1), forced.dy)
    forced.hx_ext = np.full((nx + 2, 1), forced.dx)
    forced.hy_ext = np.full((1, ny + 2), forced.dy)
    forced.wx, forced.wy = np.full((nx + 1, 1), 0.5), np.full((1, ny + 1), 0.5)
    su, sv = O.momentum_rhs(forced, u, v)
    np.testing.assert_allclose(su, ref_u, rtol=1e-12, atol=1e-12 * np.abs(ref_u).max())
    np.testing.assert_allclose(sv, ref_v, rtol=1e-12, atol=1e-12 * np.abs(ref_v).max())


# ---------------------------------------------------------------- schemes, upwind, bodies
SCHEME_RUNS = [("cavity", dict(scheme="euler"), 10, "euler"),
               ("cavity", dict(scheme="rk2"), 10, "rk2"),
               ("cyl_outlet", dict(scheme="euler"), 8, "euler"),
               ("cyl_outlet", dict(scheme="rk2"), 8, "rk2"),
               ("cavity", dict(advection="upwind"), 10, "upwind"),
               ("cyl_outlet", dict(advection="upwind"), 8, "upwind"),
               ("cyl_outlet", dict(advection="upwind", upwind_blend=0.4), 8, "upwind04")]


@pytest.mark.parametrize("name,kw,steps,label", SCHEME_RUNS,
                         ids=[f"{r[0]}-{r[3]}" for r in SCHEME_RUNS])
def test_schemes_and_upwind_match_reference(name, kw, steps, label):
    """Euler and RK2 (solver.py:699-707) and the donor-cell blend rule
    (solver.py:734-756: config blend, else min(1, 1.2 max Courant)) against
    the reference's Simulation.run."""
    g = golden("schemes")
    tag = f"{name}_{label}"
    case = oracle_case(name, pressure_solver="direct", tol=1e-12, **kw)
    u0, v0 = initial(name, case)
    u, v, p, hist, t = O.run(case, steps, u0, v0)
    scale = max(np.abs(g[f"{tag}_u"]).max(), 1.0)
    assert np.abs(u - g[f"{tag}_u"]).max() <= 1e-10 * scale
    assert np.abs(v - g[f"{tag}_v"]).max() <= 1e-10 * scale
    pscale = max(np.abs(g[f"{tag}_p"]).max(), 1.0)
    assert np.abs(p - g[f"{tag}_p"]).max() <= 1e-8 * pscale
    np.testing.assert_allclose(hist["dt"], g[f"{tag}_dt"], rtol=1e-12)
    np.testing.assert_allclose(hist["fx"], g[f"{tag}_fx"], rtol=1e-8, atol=1e-10)
    assert hist["blend"][-1] == pytest.approx(float(g[f"{tag}_blend"]), rel=1e-12)
    assert max(hist["max_div"]) < 1e-9


def test_schemes_differ_from_each_other():
    """The three schemes are distinct integrators: a restatement that ran RK3
    under every name would pass the comparisons above only if the fixtures
    were identical, which they are not."""
    g = golden("schemes")
    e, r2 = g["cavity_euler_u"], g["cavity_rk2_u"]
    r3 = golden("trajectory")["cavity_direct_u"]
    assert np.abs(e - r2).max() > 1e-6 and np.abs(r2 - r3).max() > 0
    assert float(g["cavity_upwind_blend"]) > 0.0


def tandem_case(**kw):
    g = golden("schemes")
    lead, trail = g["tandem_lead"], g["tandem_trail"]
    args = dict(nx=96, ny=48, lx=24.0, ly=12.0, re=40.0, dt=0.02, cfl_max=0.4,
                kinds=("inlet", "pressure_outlet", "symmetry", "symmetry"),
                velocity=(1.0, 0.0, 0.0, 0.0), solid=lead | trail, bodies=(lead, trail),
                pressure_solver="direct", tol=1e-12)
    args.update(kw)
    return O.Case(**args), g


def test_tandem_bodies_forces_match_reference():
    """Two cylinders four diameters apart (test_multibody.py:43-50, 238-246):
    per-body reaction forces at every step against the reference, and their
    sum equal to the total to round-off (test_multibody.py:266-273)."""
    case, g = tandem_case()
    steps = 30
    u, v, p, hist, t = O.run(case, steps, 1.0, None)
    assert np.abs(u - g["tandem_u"]).max() <= 1e-10 * np.abs(g["tandem_u"]).max()
    assert np.abs(v - g["tandem_v"]).max() <= 1e-10 * np.abs(g["tandem_u"]).max()
    fb = np.array(hist["fbody"])
    np.testing.assert_allclose(fb, g["tandem_fbody"][:steps], rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(hist["fx"], g["tandem_fx"][:steps], rtol=1e-8, atol=1e-10)
    assert np.abs(fb[:, 0] + fb[:, 2] - np.array(hist["fx"])).max() < 1e-12
    assert np.abs(fb[:, 1] + fb[:, 3] - np.array(hist["fy"])).max() < 1e-12
    # the trailing body is not the leading one (the split carries information)
    assert abs(fb[-1, 0] - fb[-1, 2]) > 1e-3


def test_strouhal_matches_reference():
    """strouhal_number (postprocess.py:116-142) on adaptive time stamps,
    including a record too short to resolve (nan)."""
    from pycfd_amd.analysis.postprocess import strouhal_number
    g = golden("schemes")
    for k in range(4):
        t, s = g[f"st{k}_t"], g[f"st{k}_s"]
        for key, (lr, ur) in (("St", (1.0, 1.0)), ("St_scaled", (0.5, 2.0))):
            want = float(g[f"st{k}_{key}"])
            got = strouhal_number(t, s, lr, ur)
            if np.isnan(want):
                assert np.isnan(got)
            else:
                assert got == want


def test_strouhal_closed_form():
    """A pure tone on a uniform base resolves to its own frequency (to one
    spectral bin), so a dropped resample or mean removal would show."""
    from pycfd_amd.analysis.postprocess import strouhal_number
    t = np.linspace(0.0, 200.0, 4001)
    for f0 in (0.1, 0.2, 0.35):
        st = strouhal_number(t, 3.0 + np.sin(2 * np.pi * f0 * t), 1.0, 1.0)
        assert abs(st - f0) <= 1.0 / 200.0
