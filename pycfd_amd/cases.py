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
