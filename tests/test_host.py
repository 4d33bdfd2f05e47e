"""Host-side setup of the engine, checked on the CPU against the oracle and the
golden vectors: operator assembly, multigrid hierarchy, the stencil
dictionaries handed to the device, the coarsest pseudo-inverse, inlet lines
and configuration validation."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O
from pycfd_amd.config import BCKind, BCSpec, SimulationConfig
from pycfd_amd.core.boundary import BoundaryManager
from pycfd_amd.core.mesh import StructuredMesh
from pycfd_amd.core.multigrid import MultigridHierarchy, coarsest_pseudo_inverse
from pycfd_amd.core.pressure import assemble_poisson_matrix
from tests.cases_def import CASES, oracle_case, product_config
from tests.conftest import golden


def product_system(name):
    cfg, body = product_config(name)
    mesh = StructuredMesh.from_config(cfg)
    bm = BoundaryManager(cfg.boundary_config, mesh)
    return assemble_poisson_matrix(mesh, bm.periodic_x, bm.periodic_y,
                                   body.mask if body else None,
                                   bm.dirichlet_pressure() or None)


@pytest.mark.parametrize("name", list(CASES))
def test_assembly_equals_oracle(name):
    S = product_system(name)
    R = O.assemble(oracle_case(name))
    assert abs(S.A - R.A).max() == 0.0
    assert np.array_equal(S.fluid, R.fluid)
    assert np.array_equal(S.bc_rhs, R.bc_rhs)
    assert S.singular == R.singular and S.pin == R.pin


@pytest.mark.parametrize("name", ["cavity", "periodic", "cyl_outlet", "cyl_pout"])
@pytest.mark.parametrize("sweeps", [1, 2])
def test_hierarchy_equals_reference(name, sweeps):
    g = golden("multigrid")
    tag = f"{name}_s{sweeps}"
    S = product_system(name)
    H = MultigridHierarchy(S.A, S.shape, S.fluid, S.singular, sweeps, sweeps,
                           (S.periodic_x, S.periodic_y))
    assert H.unknowns == list(g[f"{tag}_unknowns"])
    assert [lv.A.nnz for lv in H.levels] == list(g[f"{tag}_nnz"])
    np.testing.assert_allclose([lv.weight for lv in H.levels], g[f"{tag}_weights"], rtol=1e-13)


@pytest.mark.parametrize("name,wide", [("cavity", False), ("periodic", False),
                                       ("cyl_outlet", False), ("cyl_pout", False),
                                       ("s_cyl", False), ("s_pchannel", False),
                                       ("s_cyl", True)])
def test_stencil_dictionary_rebuilds_operator(name, wide):
    """cls/table/offsets reproduce every coarse operator exactly, with the
    16-bit class index or the 32-bit one of wide levels."""
    S = product_system(name)
    H = MultigridHierarchy(S.A, S.shape, S.fluid, S.singular, 1, 1,
                           (S.periodic_x, S.periodic_y))
    for lv, d in zip(H.levels[1:], H.device_levels(wide_classes=wide)[1:]):
        assert d["cls"].dtype == (np.uint32 if wide else np.uint16)
        nx, ny = lv.shape
        n = nx * ny
        rows, cols, vals = [], [], []
        for k in range(n):
            row = d["table"][d["cls"][k]]
            i, j = divmod(k, ny)
            for s in np.flatnonzero(row):
                ii, jj = i + int(d["di"][s]), j + int(d["dj"][s])
                if S.periodic_x:
                    ii %= nx
                if S.periodic_y:
                    jj %= ny
                assert 0 <= ii < nx and 0 <= jj < ny
                rows.append(k)
                cols.append(ii * ny + jj)
                vals.append(row[s])
        B = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
        assert abs(B - lv.A).max() == 0.0
        act = d["class_active"][d["cls"]].astype(bool)
        assert np.array_equal(act, lv.active)
        np.testing.assert_array_equal(d["class_dinv"][d["cls"]], lv.dinv)


@pytest.mark.parametrize("name", ["cavity", "cyl_outlet", "cyl_pout"])
def test_coarsest_pseudo_inverse_equals_pinned_solve(name):
    S = product_system(name)
    H = MultigridHierarchy(S.A, S.shape, S.fluid, S.singular, 1, 1,
                           (S.periodic_x, S.periodic_y))
    R = O.Hierarchy(O.assemble(oracle_case(name)).A, S.shape, S.fluid, S.singular, 1)
    C = coarsest_pseudo_inverse(H.levels[-1])
    rng = np.random.default_rng(11)
    for _ in range(3):
        b = rng.standard_normal(C.shape[0])
        np.testing.assert_allclose(C @ b, R._coarsest(R.levels[-1], b), rtol=1e-10,
                                   atol=1e-12 * np.abs(C).max() * np.abs(b).max())


@pytest.mark.parametrize("name", ["cavity", "periodic", "cyl_outlet", "cyl_pout"])
@pytest.mark.parametrize("coarse_max", [40, 150, 2048])
def test_truncated_device_levels_match_oracle(name, coarse_max):
    """device_levels(coarse_max) stops at the first level with at most
    coarse_max unknowns; its dense operator equals the oracle's exact solve
    on that level."""
    S = product_system(name)
    H = MultigridHierarchy(S.A, S.shape, S.fluid, S.singular, 1, 1,
                           (S.periodic_x, S.periodic_y))
    R = O.Hierarchy(O.assemble(oracle_case(name)).A, S.shape, S.fluid, S.singular, 1,
                    coarse_max=coarse_max)
    d = H.device_levels(coarse_max)
    assert [L["nx"] * L["ny"] for L in d] == [lv["A"].shape[0] for lv in R.levels]
    assert (d[-1]["fx"], d[-1]["fy"]) == (1, 1)
    C = d[-1]["coarse_inv"]
    rng = np.random.default_rng(12)
    for _ in range(3):
        b = rng.standard_normal(C.shape[0])
        np.testing.assert_allclose(C @ b, R._coarsest(R.levels[-1], b), rtol=1e-10,
                                   atol=1e-12 * np.abs(C).max() * np.abs(b).max())


def test_inlet_lines_match_oracle():
    cfg, _ = product_config("channel")
    mesh = StructuredMesh.from_config(cfg)
    bm = BoundaryManager(cfg.boundary_config, mesh)
    line = bm.conditions["left"].inlet_line(mesh)
    case = oracle_case("channel")
    ref = O.inlet_value(case, 0, 0, -1, case.ny + 2)
    assert np.array_equal(line, ref)


def test_config_validation_messages():
    with pytest.raises(ValueError, match="at least 4x4"):
        SimulationConfig(nx=3)
    with pytest.raises(ValueError, match="mg_sweeps"):
        SimulationConfig(mg_sweeps=0)
    with pytest.raises(ValueError, match="periodicity"):
        SimulationConfig(boundary_config={"left": BCSpec(BCKind.PERIODIC),
                                          "right": BCSpec(BCKind.NO_SLIP),
                                          "bottom": BCSpec(), "top": BCSpec()})
    with pytest.raises(ValueError, match="precision"):
        SimulationConfig(precision="fp16")
    cfg = SimulationConfig(pressure_solver="mgcg")
    assert SimulationConfig.from_dict(cfg.to_dict()) == cfg


def test_sor_rejects_odd_periodic_colouring():
    from pycfd_amd.core.pressure import SORSolver
    mesh = StructuredMesh(5, 6)
    S = assemble_poisson_matrix(mesh, True, False)
    with pytest.raises(ValueError, match="red/black"):
        SORSolver(S)


# ---------------------------------------------------------------- checkpoints
def test_checkpoint_round_trip_is_exact(tmp_path):
    """test_validation.py:321-336 on host arrays: fields, clock and config come
    back exactly; the archive keeps the reference's keys."""
    from pycfd_amd.analysis.export import (checkpoint_provenance, load_checkpoint,
                                           load_checkpoint_with_guess, save_checkpoint)
    from pycfd_amd.config import BCKind, BCSpec, SimulationConfig
    from pycfd_amd.core.fields import FlowField
    from pycfd_amd.core.mesh import StructuredMesh
    bcs = {w: BCSpec(BCKind.NO_SLIP) for w in ("left", "right", "bottom")}
    bcs["top"] = BCSpec(BCKind.MOVING_WALL, velocity=1.0)
    cfg = SimulationConfig(nx=12, ny=10, re=250, dt=3e-3, boundary_config=bcs)
    mesh = StructuredMesh.from_config(cfg)
    rng = np.random.default_rng(0)
    fields = FlowField(mesh, rng.standard_normal(mesh.u_shape), rng.standard_normal(mesh.v_shape),
                       rng.standard_normal(mesh.p_shape), 0.123, 7)
    path = save_checkpoint(fields, cfg, tmp_path / "ck")
    assert path.suffix == ".npz"
    with np.load(path, allow_pickle=False) as npz:
        assert {"u", "v", "p", "t", "step", "config", "provenance"} <= set(npz.files)
        assert "p_guess" not in npz.files
    restored, restored_cfg = load_checkpoint(path)
    assert np.array_equal(restored.u, fields.u)
    assert np.array_equal(restored.v, fields.v)
    assert np.array_equal(restored.p, fields.p)
    assert restored.t == fields.t and restored.step == fields.step
    assert restored_cfg == cfg
    assert restored_cfg.boundary_config["top"].velocity == 1.0
    assert checkpoint_provenance(path)["generator"] == "pycfd_amd"
    assert load_checkpoint_with_guess(path)[2] is None


def test_checkpoint_written_by_the_reference_loads_and_units_match(tmp_path):
    """A checkpoint the reference wrote (export.py:214-235) loads here, and a
    checkpoint written here carries the reference's units block, also under a
    provenance record the caller supplied (export.py:69-84)."""
    import json
    from pycfd_amd.analysis.export import (checkpoint_provenance, load_checkpoint,
                                           save_checkpoint)
    from tests.conftest import GOLDEN
    ref = GOLDEN / "ref_ckpt.npz"
    fields, cfg = load_checkpoint(ref)
    with np.load(ref, allow_pickle=False) as npz:
        assert np.array_equal(fields.u, npz["u"]) and np.array_equal(fields.p, npz["p"])
        assert fields.step == int(npz["step"]) == 3
        ref_units = json.loads(str(npz["provenance"]))["units"]
    assert (cfg.nx, cfg.ny) == (24, 20)
    ours = save_checkpoint(fields, cfg, tmp_path / "ours.npz")
    assert checkpoint_provenance(ours)["units"] == ref_units
    sidecar = json.loads(ours.with_suffix(".provenance.json").read_text())
    assert sidecar["units"] == ref_units
    custom = save_checkpoint(fields, cfg, tmp_path / "custom.npz",
                             provenance={"study": "supplied by the caller"})
    with np.load(GOLDEN / "ref_ckpt_custom.npz", allow_pickle=False) as npz:
        want = json.loads(str(npz["provenance"]))
    assert checkpoint_provenance(custom) == want


def test_checkpoint_rejects_mismatched_shapes_and_missing_files(tmp_path):
    from pycfd_amd.analysis.export import load_checkpoint, save_checkpoint
    from pycfd_amd.config import SimulationConfig
    from pycfd_amd.core.fields import FlowField
    from pycfd_amd.core.mesh import StructuredMesh
    with pytest.raises(FileNotFoundError):
        load_checkpoint(tmp_path / "nope.npz")
    cfg = SimulationConfig(nx=8, ny=6)
    mesh = StructuredMesh.from_config(cfg)
    fields = FlowField(mesh, np.zeros(mesh.u_shape), np.zeros(mesh.v_shape),
                       np.zeros(mesh.p_shape))
    path = save_checkpoint(fields, cfg.replace(nx=9), tmp_path / "bad.npz")
    with pytest.raises(ValueError, match="does not match"):
        load_checkpoint(path)


# ---------------------------------------------------------------- stretched meshes
@pytest.mark.parametrize("name", ["s_cavity", "s_channel", "s_cyl", "s_cylp", "s_pchannel"])
def test_stretched_assembly_equals_oracle(name):
    """pressure.py:193-207: per-face coefficients, area row scale and
    Dirichlet terms, bit for bit with the oracle (itself pinned to the
    reference's vectors); the operator is symmetric."""
    S = product_system(name)
    R = O.assemble(oracle_case(name))
    assert abs(S.A - R.A).max() == 0.0
    assert np.array_equal(S.fluid, R.fluid)
    assert np.array_equal(S.bc_rhs, R.bc_rhs)
    assert np.array_equal(np.asarray(S.row_scale), np.asarray(R.row_scale))
    assert S.singular == R.singular and S.pin == R.pin
    assert abs(S.A - S.A.T).max() <= 1e-12 * abs(S.A).max()


@pytest.mark.parametrize("name", ["s_cavity", "s_cyl"])
def test_stretched_metrics_match_oracle(name):
    """mesh.py:160-211 metric arrays, bit for bit."""
    cfg, _ = product_config(name)
    m = StructuredMesh.from_config(cfg).metrics
    c = oracle_case(name)
    for a, b in ((m.hx, c.hx), (m.hy, c.hy), (m.hxu, c.hxu), (m.hyv, c.hyv),
                 (m.hx_ext, c.hx_ext), (m.hy_ext, c.hy_ext), (m.wx_corner, c.wx),
                 (m.wy_corner, c.wy), (m.cell_areas, c.cell_areas)):
        assert np.array_equal(a, b)
    assert (m.min_hx, m.min_hy) == (c.min_hx, c.min_hy)


def test_stretched_periodic_axis_is_refused():
    """solver.py:209-230: a stretched axis may not be periodic."""
    from pycfd_amd.core.solver import ProjectionSolver
    cfg, _ = product_config("pchannel", stretch_x=1.05)
    with pytest.raises(ValueError, match="periodic and stretched"):
        ProjectionSolver(cfg)


# ---------------------------------------------------------------- batched cavities
def test_cavity_batch_configs_are_seeded_and_in_range():
    """The batched-cavity workload: sizes drawn from the stated set, Re
    log-uniform in [100, 1000], reproducible from the seed; each instance is
    lid_driven_cavity.build's configuration (lid_driven_cavity.py:59-85)."""
    from pycfd_amd.batch import CAVITY_SIZES, cavity_batch_configs
    a = cavity_batch_configs(96, seed=7)
    b = cavity_batch_configs(96, seed=7)
    c = cavity_batch_configs(96, seed=8)
    assert [(x.nx, x.re) for x in a] == [(x.nx, x.re) for x in b]
    assert [(x.nx, x.re) for x in a] != [(x.nx, x.re) for x in c]
    assert {x.nx for x in a} <= set(CAVITY_SIZES) and len({x.nx for x in a}) == len(CAVITY_SIZES)
    res = np.array([x.re for x in a])
    assert res.min() >= 100.0 and res.max() <= 1000.0
    # log-uniform: the median sits near sqrt(100 * 1000)
    assert 200.0 < np.median(res) < 500.0
    x = a[0]
    assert x.nx == x.ny and x.dt == 0.01 and x.cfl_max == 0.4 and x.lx == x.ly == 1.0
    assert x.boundary_config["top"].kind.value == "moving_wall"
    assert x.pressure_solver.value == "direct" and not x.use_les


def test_identical_operators_share_one_hierarchy():
    """Solvers of one operator get the same hierarchy object (built once); a
    different operator, sweep count or a freed hierarchy gets its own build,
    equal to a fresh one."""
    import gc

    from pycfd_amd.core.multigrid import shared_hierarchy
    mesh = StructuredMesh(48, 40, 1.0, 1.0)
    s1 = assemble_poisson_matrix(mesh)
    s2 = assemble_poisson_matrix(StructuredMesh(48, 40, 1.0, 1.0))
    h1 = shared_hierarchy(s1.A, s1.shape, s1.fluid, s1.singular)
    h2 = shared_hierarchy(s2.A, s2.shape, s2.fluid, s2.singular)
    assert h1 is h2
    assert shared_hierarchy(s1.A, s1.shape, s1.fluid, s1.singular, sweeps=2) is not h1
    s3 = assemble_poisson_matrix(mesh, dirichlet={"right": 0.0})
    h3 = shared_hierarchy(s3.A, s3.shape, s3.fluid, s3.singular)
    assert h3 is not h1
    fresh = MultigridHierarchy(s1.A, s1.shape, s1.fluid, s1.singular)
    assert fresh.unknowns == h1.unknowns
    for a, b in zip(fresh.levels, h1.levels):
        assert (a.A != b.A).nnz == 0 and a.weight == b.weight
    del h1, h2
    gc.collect()  # no solver holds it any more: the next request builds it again
    h4 = shared_hierarchy(s1.A, s1.shape, s1.fluid, s1.singular)
    assert h4.unknowns == fresh.unknowns
    assert [lv.weight for lv in h4.levels] == [lv.weight for lv in fresh.levels]


# --- the native hierarchy build (csrc/hierarchy.hip) ---------------------

def _native_levels(S, threads=4, coarse_max=96, every_row=0):
    """Every level of a native build: operator arrays, weight, dictionary."""
    import ctypes as C

    from pycfd_amd import _lib
    from pycfd_amd.core.multigrid import _Built, _start_vector
    nx, ny = S.shape
    A = sp.csr_matrix(S.A)
    keep = [np.ascontiguousarray(A.indptr, dtype=np.int64),
            np.ascontiguousarray(A.indices, dtype=np.int32),
            np.ascontiguousarray(A.data, dtype=np.float64),
            np.ascontiguousarray(S.fluid, dtype=np.uint8), _start_vector(nx * ny)]
    op = _lib.MGOperator()
    op.nx, op.ny = nx, ny
    op.indptr = keep[0].ctypes.data_as(C.POINTER(C.c_int64))
    op.indices = keep[1].ctypes.data_as(C.POINTER(C.c_int32))
    op.data = keep[2].ctypes.data_as(C.POINTER(C.c_double))
    op.active = keep[3].ctypes.data_as(C.POINTER(C.c_uint8))
    op.x0 = keep[4].ctypes.data_as(C.POINTER(C.c_double))
    op.singular = int(S.singular)
    op.periodic_x, op.periodic_y = int(S.periodic_x), int(S.periodic_y)
    op.threads, op.coarse_max, op.every_row = threads, coarse_max, every_row
    h = C.c_void_p(0)
    _lib.check(_lib.load().pcfd_mg_build(C.byref(op), C.byref(h)), "pcfd_mg_build")
    built = _Built(h)
    out = []
    for l in range(built.info(0).n_levels):
        d = built.copy(l, "dinv", "cls", "table", "class_dinv",
                       *(("indptr", "indices", "data") if l else ()))
        d["weight"] = built.info(l).weight
        out.append(d)
    return out


def _cylinder_system(nx, ny):
    from pycfd_amd import cases
    cfg, body, _, _ = cases.cylinder(nx=nx, ny=ny)
    mesh = StructuredMesh.from_config(cfg)
    bm = BoundaryManager(cfg.boundary_config, mesh)
    return assemble_poisson_matrix(mesh, bm.periodic_x, bm.periodic_y, body.mask,
                                   bm.dirichlet_pressure() or None)


@pytest.mark.parametrize("name", ["cavity", "periodic", "channel", "cyl_pout", "s_cyl",
                                  "s_pchannel", "cylinder256x128"])
def test_native_translated_rows_equal_every_row(name):
    """Forming each distinct coarse row once and translating it gives the
    operator that forming every row gives, bit for bit; so does any thread
    count."""
    S = _cylinder_system(256, 128) if name == "cylinder256x128" else product_system(name)
    a = _native_levels(S, threads=4, every_row=1)
    b = _native_levels(S, threads=4, every_row=0)
    c = _native_levels(S, threads=1, every_row=0)
    assert len(a) == len(b) == len(c)
    for x, y, z in zip(a, b, c):
        assert x["weight"] == y["weight"] == z["weight"]
        for k in x:
            assert np.array_equal(x[k], y[k]) and np.array_equal(y[k], z[k]), k


@pytest.mark.parametrize("name", ["cavity", "periodic", "channel", "cyl_outlet", "cyl_pout",
                                  "s_cyl", "s_pchannel"])
def test_native_hierarchy_matches_oracle(name):
    """Against the oracle's SciPy build of the reference algorithm: the same
    grids, factors and sparsity, weights and operators equal to round-off."""
    S = product_system(name)
    R = O.Hierarchy(O.assemble(oracle_case(name)).A, S.shape, S.fluid, S.singular, 1)
    H = MultigridHierarchy(S.A, S.shape, S.fluid, S.singular, 1, 1,
                           (S.periodic_x, S.periodic_y))
    assert H.unknowns == [lv["A"].shape[0] for lv in R.levels]
    for lv, rl in zip(H.levels, R.levels):
        assert lv.shape == rl["shape"]
        assert lv.A.nnz == rl["A"].nnz
        assert lv.weight == pytest.approx(rl["weight"], rel=1e-13)
        scale = abs(rl["A"]).max()
        assert abs(lv.A - rl["A"]).max() <= 1e-13 * scale
        np.testing.assert_array_equal(lv.active, rl["active"])


@pytest.mark.parametrize("name", ["cavity", "periodic", "cyl_pout", "s_cyl"])
def test_native_coarse_operators_keep_symmetry_and_null_space(name):
    """Galerkin operators of a symmetric operator are symmetric; where the
    fine operator is singular (row sums zero on the active set) every coarse
    operator keeps the constant null space on its active cells."""
    S = product_system(name)
    H = MultigridHierarchy(S.A, S.shape, S.fluid, S.singular, 1, 1,
                           (S.periodic_x, S.periodic_y))
    for lv in H.levels[1:]:
        A = lv.A
        scale = abs(A).max()
        assert abs(A - A.T).max() <= 1e-12 * scale
        if S.singular:
            rows = np.asarray(A @ lv.active.astype(float))
            assert np.abs(rows[lv.active]).max() <= 1e-12 * scale


def test_native_build_stops_at_coarse_max():
    """A build stopped at the first level with at most 2048 unknowns holds the
    same levels as the whole hierarchy down to there, bit for bit, and a
    solver asking for a coarser device level gets the hierarchy deepened."""
    S = _cylinder_system(256, 128)
    full = _native_levels(S, coarse_max=96)
    cut = _native_levels(S, coarse_max=2048)
    assert len(cut) < len(full)
    assert cut[-1]["dinv"].size <= 2048 < cut[-2]["dinv"].size
    for x, y in zip(cut, full):
        assert x["weight"] == y["weight"]
        for k in x:
            assert np.array_equal(x[k], y[k]), k
    H = MultigridHierarchy(S.A, S.shape, S.fluid, S.singular, coarse_max=2048)
    assert H.unknowns == [d["dinv"].size for d in cut]
    d = H.device_levels(96)
    assert [L["nx"] * L["ny"] for L in d] == [x["dinv"].size for x in full]


def test_native_build_rejects_bad_operators():
    from pycfd_amd.core.multigrid import MultigridHierarchy as MH
    S = product_system("cavity")
    bad = S.A.copy().tocsr()
    bad.indices[3] = S.A.shape[0] + 5
    with pytest.raises(ValueError, match="out of range"):
        MH(bad, S.shape, S.fluid, S.singular)
    with pytest.raises(ValueError, match="do not fit"):
        MH(S.A, (S.shape[0] + 1, S.shape[1]), S.fluid, S.singular)


@pytest.mark.parametrize("nx,ny,lx,ly", [(45, 31, 1.0, 1.0),     # odd extents: singleton blocks
                                         (40, 24, 4.0, 0.5),     # strongly anisotropic cells
                                         (33, 96, 1.0, 1.0)])
def test_native_hierarchy_odd_and_anisotropic_grids(nx, ny, lx, ly):
    """Odd extents (a last block of one cell on every level) and cells with an
    aspect ratio past the semi-coarsening threshold: the native build gives
    the oracle's grids, factors, sparsity and weights."""
    mesh = StructuredMesh(nx, ny, lx, ly)
    S = assemble_poisson_matrix(mesh)
    R = O.Hierarchy(O.assemble(O.Case(nx=nx, ny=ny, lx=lx, ly=ly)).A, S.shape, S.fluid,
                    S.singular, 1)
    H = MultigridHierarchy(S.A, S.shape, S.fluid, S.singular)
    assert H.unknowns == [lv["A"].shape[0] for lv in R.levels]
    for lv, rl in zip(H.levels, R.levels):
        assert lv.shape == rl["shape"] and lv.A.nnz == rl["A"].nnz
        assert lv.weight == pytest.approx(rl["weight"], rel=1e-13)
        assert abs(lv.A - rl["A"]).max() <= 1e-13 * abs(rl["A"]).max()
    factors = [lv.factors for lv in H.levels[:-1]]
    ratio = (lx / nx) ** 2 / (ly / ny) ** 2  # y strength over x strength on the fine grid
    if ratio > 2.5 or ratio < 0.4:
        assert (2, 1) in factors or (1, 2) in factors  # semi-coarsening taken
