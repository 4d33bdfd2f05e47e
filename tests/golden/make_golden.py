"""Write the golden fixtures for the projection-step parity tests.

Runs the reference implementation (kiliancmd/pycfd, pure NumPy/SciPy) from
/root/reference in the build container and records small input/output
vectors for the functions nothing else pins (momentum RHS on random fields,
short trajectories of whole steps, boundary passes, forces, multigrid
hierarchy data and V-cycle applications, solver outputs).  The reference never
travels with the repository; only these .npz files do.

Usage:  python tests/golden/make_golden.py   (re-record deliberately)
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent
sys.path.insert(0, str(REF))

from pycfd.config import BCKind, BCSpec, SimulationConfig  # noqa: E402
from pycfd.core import pressure as ref_pressure  # noqa: E402
from pycfd.core.solver import ProjectionSolver  # noqa: E402
from pycfd.core.timestepper import TimeStepper  # noqa: E402
from pycfd.geometry.obstacles import circle_mask  # noqa: E402
from pycfd.physics.incompressible import Simulation  # noqa: E402


def walls(**kinds):
    out = {w: BCSpec(BCKind.NO_SLIP) for w in ("left", "right", "bottom", "top")}
    out.update(kinds)
    return out


CASES = {
    "cavity": dict(nx=24, ny=20, lx=1.0, ly=1.0, re=100.0, dt=0.01, cfl_max=0.5,
                   boundary_config=walls(top=BCSpec(BCKind.MOVING_WALL, velocity=1.0))),
    "periodic": dict(nx=24, ny=16, lx=2 * np.pi, ly=2 * np.pi, re=50.0, dt=2e-3,
                     adaptive_dt=False,
                     boundary_config={w: BCSpec(BCKind.PERIODIC)
                                      for w in ("left", "right", "bottom", "top")}),
    "channel": dict(nx=32, ny=16, lx=4.0, ly=1.0, re=50.0, dt=0.01, cfl_max=0.4,
                    boundary_config={"left": BCSpec(BCKind.INLET, velocity=1.0,
                                                    profile="parabolic"),
                                     "right": BCSpec(BCKind.OUTLET),
                                     "bottom": BCSpec(BCKind.NO_SLIP),
                                     "top": BCSpec(BCKind.NO_SLIP)}),
    "cyl_outlet": dict(nx=48, ny=24, lx=12.0, ly=6.0, re=100.0, dt=0.02, cfl_max=0.4,
                       boundary_config={"left": BCSpec(BCKind.INLET, velocity=1.0),
                                        "right": BCSpec(BCKind.OUTLET),
                                        "bottom": BCSpec(BCKind.SYMMETRY),
                                        "top": BCSpec(BCKind.SYMMETRY)}),
    "cyl_pout": dict(nx=48, ny=24, lx=12.0, ly=6.0, re=100.0, dt=0.02, cfl_max=0.4,
                     boundary_config={"left": BCSpec(BCKind.INLET, velocity=1.0),
                                      "right": BCSpec(BCKind.PRESSURE_OUTLET, p_ref=0.0),
                                      "bottom": BCSpec(BCKind.SYMMETRY),
                                      "top": BCSpec(BCKind.SYMMETRY)}),
}


def config(name, **kw):
    base = dict(CASES[name])
    base.update(kw)
    base.setdefault("name", name)
    base.setdefault("t_end", 100.0)
    return SimulationConfig(**base)


def body_for(name, mesh):
    if name.startswith("cyl"):
        return circle_mask(mesh, (3.0, 3.0), 0.5)
    return None


def init_for(name, mesh):
    if name.startswith("cyl") or name == "channel":
        Xv, Yv = mesh.v_grid()
        return 1.0 if name.startswith("cyl") else 0.0, 0.05 * np.exp(-((Xv - 4.0) ** 2 + (Yv - 3.0) ** 2))
    if name == "periodic":
        Xu, Yu = mesh.u_grid()
        Xv, Yv = mesh.v_grid()
        return np.sin(Xu) * np.cos(Yu), -np.cos(Xv) * np.sin(Yv)
    return None, None


def rhs_fixture():
    rng = np.random.default_rng(1234)
    out = {}
    for name in ("cavity", "periodic"):
        solver = ProjectionSolver(config(name, use_numba=False))
        nx, ny = solver.mesh.shape
        u = rng.standard_normal((nx + 3, ny + 2))
        v = rng.standard_normal((nx + 2, ny + 3))
        for blend in (0.0, 0.7):
            ru, rv = solver.momentum_rhs(u, v, blend)
            tag = f"{name}_b{int(blend * 10)}"
            out[f"{tag}_u"], out[f"{tag}_v"] = u, v
            out[f"{tag}_rhs_u"], out[f"{tag}_rhs_v"] = ru.copy(), rv.copy()
        out[f"{name}_div"] = solver.divergence(u, v)
    np.savez_compressed(OUT / "momentum_rhs.npz", **out)


def boundary_fixture():
    """apply_velocity / mass balance / mask on random fields, per case."""
    rng = np.random.default_rng(99)
    out = {}
    for name in CASES:
        cfg = config(name)
        from pycfd.core.mesh import StructuredMesh
        mesh = StructuredMesh.from_config(cfg)
        body = body_for(name, mesh)
        solver = ProjectionSolver(cfg, mesh, body.mask if body else None)
        nx, ny = mesh.shape
        from pycfd.core.fields import FlowField
        u = rng.standard_normal((nx + 3, ny + 2)) + 0.5
        v = rng.standard_normal((nx + 2, ny + 3))
        out[f"{name}_u"], out[f"{name}_v"] = u.copy(), v.copy()
        for pred in (True, False):
            f = FlowField(mesh, u.copy(), v.copy(), mesh.zeros_p())
            solver.boundaries.apply_velocity(f, predictor=pred)
            out[f"{name}_bc{int(pred)}_u"], out[f"{name}_bc{int(pred)}_v"] = f.u, f.v
        f = FlowField(mesh, u.copy(), v.copy(), mesh.zeros_p())
        solver.boundaries.apply_velocity(f, predictor=True)
        solver.boundaries.enforce_global_mass_balance(f)
        solver._mask_obstacle(f.u, f.v, dt=0.01, reset=True)
        solver._wrap_periodic_faces(f.u, f.v)
        out[f"{name}_pass_u"], out[f"{name}_pass_v"] = f.u, f.v
        out[f"{name}_pass_force"] = np.array(solver.body_force_reaction)
        p = rng.standard_normal((nx + 2, ny + 2))
        out[f"{name}_p_in"] = p.copy()
        f = FlowField(mesh, u.copy(), v.copy(), p)
        solver.boundaries.apply_pressure(f)
        out[f"{name}_p_bc"] = f.p
    np.savez_compressed(OUT / "boundary.npz", **out)


def trajectory_fixture():
    """Whole steps through Simulation.run with tight solves."""
    out = {}
    runs = [("cavity", "direct", 20), ("cavity", "cg", 5), ("periodic", "direct", 10),
            ("channel", "direct", 10), ("cyl_outlet", "direct", 10),
            ("cyl_pout", "direct", 10)]
    for name, ps, steps in runs:
        cfg = config(name, pressure_solver=ps, poisson_tol=1e-12, use_numba=False)
        from pycfd.core.mesh import StructuredMesh
        mesh = StructuredMesh.from_config(cfg)
        body = body_for(name, mesh)
        u0, v0 = init_for(name, mesh)
        sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
        tag = f"{name}_{ps}"
        out[f"{tag}_u0"], out[f"{tag}_v0"], out[f"{tag}_p0"] = (
            sim.fields.u.copy(), sim.fields.v.copy(), sim.fields.p.copy())
        res = sim.run(max_steps=steps)
        h = res.history
        out[f"{tag}_u"], out[f"{tag}_v"], out[f"{tag}_p"] = sim.fields.u, sim.fields.v, sim.fields.p
        for k in ("dt", "max_div", "kinetic_energy", "fx", "fy"):
            out[f"{tag}_{k}"] = np.asarray(h[k])
        out[f"{tag}_t"] = np.array(sim.fields.t)
    np.savez_compressed(OUT / "trajectory.npz", **out)


def multigrid_fixture():
    from pycfd.core.mesh import StructuredMesh
    out = {}
    rng = np.random.default_rng(7)
    for name in ("cavity", "periodic", "cyl_outlet", "cyl_pout"):
        for sweeps in (1, 2):
            cfg = config(name, pressure_solver="mgcg", mg_sweeps=sweeps)
            mesh = StructuredMesh.from_config(cfg)
            body = body_for(name, mesh)
            solver = ProjectionSolver(cfg, mesh, body.mask if body else None)
            h = solver.pressure_solver.hierarchy
            tag = f"{name}_s{sweeps}"
            out[f"{tag}_unknowns"] = np.array(h.unknowns)
            out[f"{tag}_weights"] = np.array([lv.weight for lv in h.levels])
            out[f"{tag}_nnz"] = np.array([lv.A.nnz for lv in h.levels])
            r = rng.standard_normal(h.levels[0].n)
            out[f"{tag}_r"] = r
            out[f"{tag}_z"] = h.apply(r)
            rhs = rng.standard_normal(mesh.shape)
            out[f"{tag}_rhs"] = rhs
            for ps in ("direct", "cg", "sor", "multigrid", "mgcg", "jacobi"):
                if ps == "jacobi" and name != "cavity":
                    continue
                s = ref_pressure.make_pressure_solver(
                    ps, solver.poisson, tol=1e-10, maxiter=200000, mg_sweeps=sweeps)
                p = s.solve(rhs)
                out[f"{tag}_{ps}_p"] = p
                out[f"{tag}_{ps}_it"] = np.array(s.last_iterations)
                out[f"{tag}_{ps}_res"] = np.array(s.last_residual)
    np.savez_compressed(OUT / "multigrid.npz", **out)


def les_fixture():
    """Smagorinsky closure: eddy viscosity and the stress-form momentum RHS on
    random fields, and a short LES trajectory (turbulence.py:63-113,
    solver.py:467-529)."""
    from pycfd.core.mesh import StructuredMesh
    rng = np.random.default_rng(4321)
    out = {}
    for name in ("cavity", "periodic"):
        solver = ProjectionSolver(config(name, use_les=True, smagorinsky_cs=0.17))
        nx, ny = solver.mesh.shape
        u = rng.standard_normal((nx + 3, ny + 2))
        v = rng.standard_normal((nx + 2, ny + 3))
        nu_c, nu_k = solver.turbulence.eddy_viscosity(u, v, solver.nu)
        out[f"{name}_u"], out[f"{name}_v"] = u, v
        out[f"{name}_nu_c"], out[f"{name}_nu_corner"] = nu_c, nu_k
        for blend in (0.0, 0.7):
            ru, rv = solver.momentum_rhs(u, v, blend)
            tag = f"{name}_b{int(blend * 10)}"
            out[f"{tag}_rhs_u"], out[f"{tag}_rhs_v"] = ru.copy(), rv.copy()
    for name, steps in (("cavity", 10), ("cyl_outlet", 6)):
        cfg = config(name, pressure_solver="direct", use_les=True, smagorinsky_cs=0.17)
        mesh = StructuredMesh.from_config(cfg)
        body = body_for(name, mesh)
        u0, v0 = init_for(name, mesh)
        sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
        sim.run(max_steps=steps)
        out[f"traj_{name}_u"], out[f"traj_{name}_v"] = sim.fields.u, sim.fields.v
        out[f"traj_{name}_p"] = sim.fields.p
    np.savez_compressed(OUT / "les.npz", **out)


STRETCHED = {
    "s_cavity": dict(base="cavity", stretch_x=1.06, stretch_y=1.04, cluster_x="walls",
                     cluster_y="walls"),
    "s_channel": dict(base="channel", stretch_y=1.08, cluster_y="walls"),
    "s_cyl": dict(base="cyl_outlet", stretch_x=1.03, stretch_y=1.05, cluster_x="low",
                  cluster_y="centre"),
    "s_cylp": dict(base="cyl_pout", stretch_y=1.05, cluster_y="centre"),
}


def stretched_config(name, **kw):
    spec = dict(STRETCHED[name])
    base = spec.pop("base")
    spec.update(kw)
    return config(base, **spec), base


def stretched_fixture():
    """Stretched meshes (mesh.py:122-310): momentum RHS and divergence on
    random fields, the predictor boundary pass with its force, the Poisson
    operator, and short trajectories (solver.py:334-683, pressure.py:193-207)."""
    from pycfd.core.mesh import StructuredMesh
    rng = np.random.default_rng(2024)
    out = {}
    for name in STRETCHED:
        cfg, base = stretched_config(name, use_numba=False)
        mesh = StructuredMesh.from_config(cfg)
        body = body_for(base, mesh)
        solver = ProjectionSolver(cfg, mesh, body.mask if body else None)
        nx, ny = mesh.shape
        u = rng.standard_normal((nx + 3, ny + 2))
        v = rng.standard_normal((nx + 2, ny + 3))
        out[f"{name}_u"], out[f"{name}_v"] = u, v
        out[f"{name}_dx_cells"], out[f"{name}_dy_cells"] = mesh.dx_cells, mesh.dy_cells
        for blend in (0.0, 0.7):
            ru, rv = solver.momentum_rhs(u, v, blend)
            out[f"{name}_b{int(blend * 10)}_rhs_u"] = ru.copy()
            out[f"{name}_b{int(blend * 10)}_rhs_v"] = rv.copy()
        out[f"{name}_div"] = solver.divergence(u, v)
        from pycfd.core.fields import FlowField
        us, vs = u.copy(), v.copy()
        fld = FlowField(mesh, us, vs, mesh.zeros_p())
        solver.boundaries.apply_velocity(fld, predictor=True)
        solver.boundaries.enforce_global_mass_balance(fld)
        solver._mask_obstacle(us, vs, 0.01, reset=True)
        out[f"{name}_pass_u"], out[f"{name}_pass_v"] = us, vs
        out[f"{name}_pass_force"] = np.array(solver.body_force_reaction)
        A = solver.poisson.A
        x = rng.standard_normal(A.shape[0])
        out[f"{name}_Ax_x"], out[f"{name}_Ax"] = x, A @ x
        out[f"{name}_row_scale"] = np.asarray(solver.poisson.row_scale, dtype=float)
        out[f"{name}_bc_rhs"] = solver.poisson.bc_rhs
        steps = 8
        cfg2, _ = stretched_config(name, pressure_solver="direct", poisson_tol=1e-12,
                                   use_numba=False)
        u0, v0 = init_for(base, mesh)
        sim = Simulation(cfg2, obstacle=body, u_init=u0, v_init=v0)
        res = sim.run(max_steps=steps)
        out[f"traj_{name}_u"], out[f"traj_{name}_v"] = sim.fields.u, sim.fields.v
        out[f"traj_{name}_p"] = sim.fields.p
        for k in ("dt", "max_div", "kinetic_energy", "fx", "fy"):
            out[f"traj_{name}_{k}"] = np.asarray(res.history[k])
    np.savez_compressed(OUT / "stretched.npz", **out)


def les_stretched_fixture():
    """Smagorinsky closure on stretched meshes: per-cell filter width and the
    array-metric strain rates and stress form (turbulence.py:40-113,
    solver.py:467-529) on random fields, and short LES trajectories."""
    from pycfd.core.mesh import StructuredMesh
    rng = np.random.default_rng(777)
    out = {}
    for name in ("s_cavity", "s_cyl"):
        cfg, base = stretched_config(name, use_les=True, smagorinsky_cs=0.17, use_numba=False)
        mesh = StructuredMesh.from_config(cfg)
        body = body_for(base, mesh)
        solver = ProjectionSolver(cfg, mesh, body.mask if body else None)
        nx, ny = mesh.shape
        u = rng.standard_normal((nx + 3, ny + 2))
        v = rng.standard_normal((nx + 2, ny + 3))
        nu_c, nu_k = solver.turbulence.eddy_viscosity(u, v, solver.nu)
        out[f"{name}_u"], out[f"{name}_v"] = u, v
        out[f"{name}_nu_c"], out[f"{name}_nu_corner"] = nu_c, nu_k
        for blend in (0.0, 0.7):
            ru, rv = solver.momentum_rhs(u, v, blend)
            tag = f"{name}_b{int(blend * 10)}"
            out[f"{tag}_rhs_u"], out[f"{tag}_rhs_v"] = ru.copy(), rv.copy()
        cfg2, _ = stretched_config(name, pressure_solver="direct", poisson_tol=1e-12,
                                   use_les=True, smagorinsky_cs=0.17, use_numba=False)
        u0, v0 = init_for(base, mesh)
        sim = Simulation(cfg2, obstacle=body, u_init=u0, v_init=v0)
        res = sim.run(max_steps=8)
        out[f"traj_{name}_u"], out[f"traj_{name}_v"] = sim.fields.u, sim.fields.v
        out[f"traj_{name}_p"] = sim.fields.p
        for k in ("dt", "kinetic_energy", "fx"):
            out[f"traj_{name}_{k}"] = np.asarray(res.history[k])
    np.savez_compressed(OUT / "les_stretched.npz", **out)


def schemes_fixture():
    """Euler and RK2 steps (solver.py:699-707), the upwind blend rule
    (solver.py:734-756), two bodies with per-body forces (solver.py:136-146,
    197-207) and the Strouhal reduction (postprocess.py:116-142)."""
    from pycfd.analysis.postprocess import strouhal_number
    from pycfd.cases import cylinder_flow
    from pycfd.core.mesh import StructuredMesh
    out = {}
    runs = [("cavity", dict(time_scheme="euler"), 10, "euler"),
            ("cavity", dict(time_scheme="rk2"), 10, "rk2"),
            ("cyl_outlet", dict(time_scheme="euler"), 8, "euler"),
            ("cyl_outlet", dict(time_scheme="rk2"), 8, "rk2"),
            ("cavity", dict(advection_scheme="upwind"), 10, "upwind"),
            ("cyl_outlet", dict(advection_scheme="upwind"), 8, "upwind"),
            ("cyl_outlet", dict(advection_scheme="upwind", upwind_blend=0.4), 8, "upwind04")]
    for name, kw, steps, label in runs:
        cfg = config(name, pressure_solver="direct", poisson_tol=1e-12, use_numba=False, **kw)
        mesh = StructuredMesh.from_config(cfg)
        body = body_for(name, mesh)
        u0, v0 = init_for(name, mesh)
        sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
        res = sim.run(max_steps=steps)
        tag = f"{name}_{label}"
        out[f"{tag}_u"], out[f"{tag}_v"], out[f"{tag}_p"] = sim.fields.u, sim.fields.v, sim.fields.p
        for k in ("dt", "max_div", "fx", "fy"):
            out[f"{tag}_{k}"] = np.asarray(res.history[k])
        out[f"{tag}_blend"] = np.array(sim.solver._blend)
    # two bodies in tandem (test_multibody.py:43-50, 238-246) on a short run
    m = StructuredMesh(96, 48, 24.0, 12.0)
    from pycfd.geometry.obstacles import ObstacleGroup
    lead = circle_mask(m, (6.0, 6.0), 0.5, name="lead")
    trail = circle_mask(m, (10.0, 6.0), 0.5, name="trail")
    sim = cylinder_flow.build(re=40.0, nx=96, ny=48, domain_length=24.0, domain_height=12.0,
                              obstacle=ObstacleGroup((lead, trail)), t_end=100.0)
    fx, fy, fb = [], [], []

    def record(fields, info):
        fx.append(sim.solver.body_force_reaction[0])
        fy.append(sim.solver.body_force_reaction[1])
        fb.append(np.asarray(sim.solver.body_force_reactions, dtype=float).ravel())

    sim.run(max_steps=30, callback=record, callback_every=1)
    out["tandem_config"] = np.array(sim.config.to_json())
    out["tandem_lead"], out["tandem_trail"] = lead.mask, trail.mask
    out["tandem_u"], out["tandem_v"], out["tandem_p"] = sim.fields.u, sim.fields.v, sim.fields.p
    out["tandem_fx"], out["tandem_fy"] = np.array(fx), np.array(fy)
    out["tandem_fbody"] = np.array(fb)
    out["tandem_cd_by_body"] = np.array([sim.force_coefficients_by_body()[k]
                                         for k in ("lead", "trail")])
    out["tandem_cd"] = np.array(sim.force_coefficients())
    # Strouhal on adaptive time stamps: a shedding-like signal with a harmonic
    # and noise, and a record too short to resolve
    rng = np.random.default_rng(515)
    for k, (n, f0) in enumerate(((400, 0.19), (1500, 0.1653), (64, 0.5), (12, 0.2))):
        t = np.cumsum(0.02 + 0.01 * rng.random(n))
        s = (np.sin(2 * np.pi * f0 * t) + 0.3 * np.sin(4 * np.pi * f0 * t + 0.7)
             + 0.05 * rng.standard_normal(n) + 1.3)
        out[f"st{k}_t"], out[f"st{k}_s"] = t, s
        out[f"st{k}_St"] = np.array(strouhal_number(t, s, 1.0, 1.0))
        out[f"st{k}_St_scaled"] = np.array(strouhal_number(t, s, 0.5, 2.0))
    np.savez_compressed(OUT / "schemes.npz", **out)


def cavity_c1_fixture():
    """The 128 x 128 lid-driven cavity exactly as the CPU-runnable benchmark
    states it: Re 100, CG to relative residual 1e-10 (and 1e-13, the tight
    pin), CFL 0.5 with the case's dt ceiling 0.01, 2000 SSP-RK3 steps
    (lid_driven_cavity.py:59-85, pressure.py:366-398)."""
    import time
    out = {}
    for tol, tag in ((1e-10, "t10"), (1e-13, "t13")):
        cfg = SimulationConfig(
            name="cavity_c1", nx=128, ny=128, lx=1.0, ly=1.0, re=100.0, u_ref=1.0, l_ref=1.0,
            dt=0.01, cfl_max=0.5, t_end=1e9, pressure_solver="cg", poisson_tol=tol,
            use_numba=False, boundary_config=walls(top=BCSpec(BCKind.MOVING_WALL, velocity=1.0)))
        sim = Simulation(cfg)
        its = []

        def record(fields, info):
            its.append(sim.solver.pressure_solver.last_iterations)

        w0 = time.perf_counter()
        res = sim.run(max_steps=2000, callback=record, callback_every=1)
        wall = time.perf_counter() - w0
        print(f"C1 tol {tol:g}: {res.steps} steps in {wall:.1f} s", flush=True)
        out[f"{tag}_u"], out[f"{tag}_v"], out[f"{tag}_p"] = sim.fields.u, sim.fields.v, sim.fields.p
        for k in ("dt", "max_div", "kinetic_energy", "residual"):
            out[f"{tag}_{k}"] = np.asarray(res.history[k])
        out[f"{tag}_iters"] = np.array(its[:-1])
        out[f"{tag}_t"] = np.array(sim.fields.t)
        out[f"{tag}_wall"] = np.array(wall)
    np.savez_compressed(OUT / "cavity_c1.npz", **out)


def checkpoint_fixture():
    """A checkpoint written by the reference's save_checkpoint
    (export.py:214-235): a 24 x 20 cavity after 3 steps, once with its own
    provenance and once with a caller-supplied record."""
    import shutil
    import tempfile
    from pycfd.analysis.export import save_checkpoint
    cfg = config("cavity", pressure_solver="direct", use_numba=False)
    sim = Simulation(cfg)
    sim.run(max_steps=3)
    tmp = Path(tempfile.mkdtemp())
    save_checkpoint(sim.fields, cfg, tmp / "ref_ckpt.npz")
    save_checkpoint(sim.fields, cfg, tmp / "ref_ckpt_custom.npz",
                    provenance={"study": "supplied by the caller"})
    for name in ("ref_ckpt.npz", "ref_ckpt_custom.npz"):
        shutil.copy(tmp / name, OUT / name)
    shutil.rmtree(tmp)


def mask_fixture():
    from pycfd.core.mesh import StructuredMesh
    out = {}
    for nx, ny, lx, ly, cx in ((2048, 1024, 32.0, 16.0, 4.0), (512, 256, 16.0, 8.0, 4.0),
                                (48, 24, 12.0, 6.0, 3.0)):
        mesh = StructuredMesh(nx, ny, lx, ly)
        ob = circle_mask(mesh, (cx, ly / 2), 0.5)
        out[f"m{nx}x{ny}"] = np.packbits(ob.mask.ravel())
    np.savez_compressed(OUT / "circle_mask.npz", **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["rhs", "boundary", "trajectory", "multigrid", "mask", "les",
                             "stretched", "les_stretched", "schemes", "cavity_c1", "checkpoint"]
    for w in which:
        print("writing", w, flush=True)
        {"rhs": rhs_fixture, "boundary": boundary_fixture, "trajectory": trajectory_fixture,
         "multigrid": multigrid_fixture, "mask": mask_fixture, "les": les_fixture,
         "stretched": stretched_fixture, "les_stretched": les_stretched_fixture,
         "schemes": schemes_fixture, "cavity_c1": cavity_c1_fixture,
         "checkpoint": checkpoint_fixture}[w]()
