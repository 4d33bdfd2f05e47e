"""Case builders for the benchmark workloads.

Each builder returns the ``SimulationConfig`` plus the obstacle and initial
condition that ``Simulation`` takes, set up as the reference's case modules
do (pycfd/cases/lid_driven_cavity.py:59-85, taylor_green.py:64-115,
cylinder_flow.py:175-284, 394-406) with the settings of the three measured
workloads:

* ``cavity``        -- lid-driven cavity, unit square, Re 100, lid u = 1.
* ``taylor_green``  -- periodic [0, 2 pi]^2, u = sin x cos y, v = -cos x sin y,
                       fixed dt, multigrid V(2,2).
* ``cylinder``      -- 32 x 16 D channel, D = 1 at (4, 8), uniform inflow,
                       zero-gradient OUTLET, slip walls, Re 100, v seeded by a
                       1e-3 Gaussian blob one diameter behind the body.
"""

from __future__ import annotations

import numpy as np

from .config import BCKind, BCSpec, SimulationConfig
from .core.mesh import StructuredMesh
from .geometry.obstacles import circle_mask


def cavity(nx: int = 128, ny: int | None = None, re: float = 100.0,
           pressure_solver: str = "cg", cfl_max: float = 0.5, dt: float = 0.01,
           **overrides):
    """Lid-driven cavity (lid_driven_cavity.py:59-85); dt ceiling 0.01."""
    ny = nx if ny is None else ny
    overrides.setdefault("name", f"cavity_Re{re:g}")
    overrides.setdefault("t_end", 1.0e9)
    cfg = SimulationConfig(
        nx=nx, ny=ny, lx=1.0, ly=1.0, re=re, u_ref=1.0, l_ref=1.0, dt=dt, cfl_max=cfl_max,
        pressure_solver=pressure_solver,
        boundary_config={
            "left": BCSpec(BCKind.NO_SLIP), "right": BCSpec(BCKind.NO_SLIP),
            "bottom": BCSpec(BCKind.NO_SLIP),
            "top": BCSpec(BCKind.MOVING_WALL, velocity=1.0),
        },
        **overrides)
    return cfg, None, None, None


def taylor_green(n: int = 1024, nu: float = 0.01, dt: float = 1.0e-3,
                 pressure_solver: str = "multigrid", mg_sweeps: int = 2, **overrides):
    """Phase-shifted Taylor-Green vortex: u = sin x cos y, v = -cos x sin y
    (taylor_green.py:64-68 shifted by (pi/2, pi/2)); exact decay exp(-2 nu t)."""
    overrides.setdefault("name", f"taylor_green_{n}")
    overrides.setdefault("t_end", 1.0e9)
    L = 2.0 * np.pi
    cfg = SimulationConfig(
        nx=n, ny=n, lx=L, ly=L, re=1.0 / nu, u_ref=1.0, l_ref=1.0, dt=dt, adaptive_dt=False,
        pressure_solver=pressure_solver, mg_sweeps=mg_sweeps,
        boundary_config={w: BCSpec(BCKind.PERIODIC) for w in ("left", "right", "bottom", "top")},
        **overrides)
    mesh = StructuredMesh.from_config(cfg)
    Xu, Yu = mesh.u_grid()
    Xv, Yv = mesh.v_grid()
    return cfg, None, np.sin(Xu) * np.cos(Yu), -np.cos(Xv) * np.sin(Yv)


def taylor_green_exact(mesh: StructuredMesh, t: float, nu: float):
    Xu, Yu = mesh.u_grid()
    Xv, Yv = mesh.v_grid()
    decay = np.exp(-2.0 * nu * t)
    return np.sin(Xu) * np.cos(Yu) * decay, -np.cos(Xv) * np.sin(Yv) * decay


def cylinder(nx: int = 2048, ny: int = 1024, re: float = 100.0,
             domain_length: float = 32.0, domain_height: float = 16.0,
             cylinder_x: float = 4.0, perturbation: float = 1.0e-3,
             pressure_solver: str = "mgcg", outlet: str = "outlet", dt: float = 0.02,
             cfl_max: float = 0.4, **overrides):
    """Flow past a cylinder (cylinder_flow.py:175-284) with a velocity outlet."""
    overrides.setdefault("name", f"cylinder_Re{re:g}")
    overrides.setdefault("t_end", 1.0e9)
    mesh = StructuredMesh(nx, ny, domain_length, domain_height)
    cy = domain_height / 2.0
    body = circle_mask(mesh, (cylinder_x, cy), 0.5, name="cylinder")
    right = (BCSpec(BCKind.OUTLET) if outlet == "outlet"
             else BCSpec(BCKind.PRESSURE_OUTLET, p_ref=0.0))
    cfg = SimulationConfig(
        nx=nx, ny=ny, lx=domain_length, ly=domain_height, re=re, u_ref=1.0,
        l_ref=body.characteristic_length, dt=dt, cfl_max=cfl_max,
        pressure_solver=pressure_solver,
        boundary_config={
            "left": BCSpec(BCKind.INLET, velocity=1.0, profile="uniform"),
            "right": right,
            "bottom": BCSpec(BCKind.SYMMETRY), "top": BCSpec(BCKind.SYMMETRY),
        },
        **overrides)
    Xv, Yv = mesh.v_grid()
    blob = np.exp(-(((Xv - (cylinder_x + 1.0)) / 1.0) ** 2 + ((Yv - cy) / 0.5) ** 2))
    return cfg, body, 1.0, perturbation * blob


def run_cylinder_record(steps: int, precision: str = "fp64", nx: int = 2048, ny: int = 1024,
                        tol: float = 2e-11, progress_every: int = 1000, warm_start: int = 0,
                        **overrides) -> dict:
    """The cylinder case run as cylinder_flow.run measures it
    (cylinder_flow.py:458-492, 590-600): Cd and Cl from the body reaction
    every 10 steps, a wake probe of v at (lx/2, ly/2 + D/4), mean Cd and Cl
    rms over t > 0.5 t_end (timeseries.py:255-259), and the Strouhal number
    of the whole probe record (postprocess.py:116-142).  ``warm_start`` is
    the solver's ``warm_start_order`` (0: the reference's start)."""
    import sys
    import time

    from .analysis.postprocess import force_coefficients, late_time_mean, strouhal_number
    from .physics.incompressible import Simulation
    cfg, body, u0, v0 = cylinder(nx=nx, ny=ny, precision=precision, poisson_tol=tol,
                                 **overrides)
    sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
    sim.solver.warm_start_order = warm_start
    probe = sim.add_probe(cfg.lx * 0.5, cfg.ly * 0.5 + 0.25, "wake")
    hist = {"t": [], "cd": [], "cl": []}
    t0 = time.perf_counter()

    def record(fields, info):
        cd, cl = force_coefficients(sim.solver.body_force_reaction, cfg.u_ref, cfg.l_ref)
        hist["t"].append(fields.t)
        hist["cd"].append(cd)
        hist["cl"].append(cl)
        if progress_every and fields.step % progress_every == 0:
            print(f"[{precision}] step {fields.step} t {fields.t:.3f} Cd {cd:.5f} Cl {cl:+.5f} "
                  f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)

    res = sim.run(max_steps=steps, callback=record, callback_every=10)
    wall = time.perf_counter() - t0
    t = np.asarray(hist["t"])
    cd, cl = np.asarray(hist["cd"]), np.asarray(hist["cl"])
    settled = t > 0.5 * t[-1]
    series = {k: np.asarray(v) for k, v in (("t", probe.t), ("v", probe.v))}
    return {
        "precision": precision, "warm_start": warm_start, "steps": res.steps, "t_end": float(sim.fields.t),
        "wall_s": round(wall, 2), "mlups": round(nx * ny * res.steps / wall / 1e6, 2),
        "cd_mean": late_time_mean(t, cd), "cl_rms": float(np.sqrt(np.mean(cl[settled] ** 2))),
        "strouhal": strouhal_number(series["t"], series["v"], cfg.l_ref, cfg.u_ref),
        "max_div": float(max(res.history["max_div"])), "records": int(t.size),
        "cd_series": cd.tolist(), "cl_series": cl.tolist(), "t_series": t.tolist(),
    }


if __name__ == "__main__":  # python -m pycfd_amd.cases cylinder-record --steps N ...
    import argparse
    import json

    ap = argparse.ArgumentParser(description="long cylinder run: Cd, Cl rms and St per path")
    ap.add_argument("case", choices=["cylinder-record"])
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--precision", default="fp64,mixed")
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--tol", type=float, default=2e-11)
    ap.add_argument("--warm-start", type=int, default=0,
                    help="extrapolation order of each solve's start (0: the last pressure)")
    ap.add_argument("--out", default=None, help="write the full records (JSON) here")
    a = ap.parse_args()
    runs = {p: run_cylinder_record(a.steps, p, a.nx, a.ny, a.tol, warm_start=a.warm_start)
            for p in a.precision.split(",")}
    summary = {p: {k: v for k, v in r.items() if not k.endswith("_series")} for p, r in runs.items()}
    if "fp64" in runs and "mixed" in runs:
        f, m = runs["fp64"], runs["mixed"]
        summary["mixed_vs_fp64"] = {
            "cd_mean_rel_diff": abs(m["cd_mean"] - f["cd_mean"]) / abs(f["cd_mean"]),
            "cl_rms_rel_diff": abs(m["cl_rms"] - f["cl_rms"]) / abs(f["cl_rms"]),
            "strouhal_diff": abs(m["strouhal"] - f["strouhal"]),
            "cd_series_max_rel_diff": float(np.max(np.abs(np.asarray(m["cd_series"])
                                                         - np.asarray(f["cd_series"]))
                                                  / np.abs(np.asarray(f["cd_series"])))),
        }
    print(json.dumps(summary, indent=1))
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(runs, fh)
