"""The benchmark workload itself against the oracle.

The 2048 x 1024 cylinder (32 x 16 D, Re 100, uniform inflow, zero-gradient
OUTLET, slip walls, D = 1 at (4, 8), v seeded with a 1e-3 Gaussian blob):
the initial projection and two SSP-RK3 steps on the GPU, FP64 and mixed,
against the oracle's own run of the same steps, both solving with
multigrid-CG to relative residual 1e-13: velocity within 1e-12 relative
(measured: about 2e-15) on both paths.  max|div u| after each step tracks the
oracle's own within 1e-12 and stays under 3e-12: the first step sits at the
FP64 floor of this grid (about eps * ||A|| * max|p|, ||A|| = 4 / dx^2 = 1.6e4),
measured 1.68e-12 in the oracle and 1.79-1.89e-12 here, the second at
5.9-6.4e-13 in both; a tighter tolerance (2e-14) does not lower it.  The
small cases pin the kernels bit for bit; this pins the full-size plan
(streaming level-0 legs, fused and row-pass stored levels, the dense
coarsest solve) on the shape the throughput is quoted on.
"""

from __future__ import annotations

import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NX, NY, LX, LY = 2048, 1024, 32.0, 16.0
TOL = 1.0e-13
STEPS = 2


def _solid():
    """Solid iff more than half of 8x8 mid-point samples lie in the disc
    (obstacles.py:87-135), as the product's circle_mask."""
    dx, dy = LX / NX, LY / NY
    offs = (np.arange(8) + 0.5) / 8
    xf = np.concatenate(([0.0], np.cumsum(np.full(NX, dx))))  # face positions as the mesh
    yf = np.concatenate(([0.0], np.cumsum(np.full(NY, dy))))
    xs = (xf[:-1, None] + offs[None, :] * dx).ravel()
    ys = (yf[:-1, None] + offs[None, :] * dy).ravel()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    return (((X - 4.0) ** 2 + (Y - 8.0) ** 2 <= 0.25).astype(float)
            .reshape(NX, 8, NY, 8).mean(axis=(1, 3)) > 0.5)


@pytest.fixture(scope="module")
def oracle_run():
    t0 = time.perf_counter()
    case = O.Case(nx=NX, ny=NY, lx=LX, ly=LY, re=100.0, dt=0.02, cfl_max=0.4,
                  kinds=("inlet", "outlet", "symmetry", "symmetry"),
                  velocity=(1.0, 0.0, 0.0, 0.0), solid=_solid(), pressure_solver="mgcg",
                  tol=TOL)
    Xv, Yv = np.meshgrid(case.xc, case.yf, indexing="ij")
    v0 = 1e-3 * np.exp(-(((Xv - 5.0) / 1.0) ** 2 + ((Yv - 8.0) / 0.5) ** 2))
    st = O.State()
    u, v, _ = O.initialize(case, st, 1.0, v0)
    divs = []
    for _ in range(STEPS):
        dt = O.compute_dt(case, u, v)
        u, v, _ = O.step(case, st, u, v, dt)
        divs.append(float(np.abs(O.divergence(case, u, v)[~case.solid]).max()))
    print(f"oracle: {NX}x{NY}, initial projection + {STEPS} steps in "
          f"{time.perf_counter() - t0:.1f} s, max|div u| per step "
          + ", ".join(f"{d:.2e}" for d in divs), flush=True)
    return case, u, v, divs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_benchmark_cylinder_matches_oracle(gpu_available, oracle_run, precision):
    from pycfd_amd import cases
    from pycfd_amd.physics.incompressible import Simulation
    case, u_ref, v_ref, div_ref = oracle_run
    cfg, body, u0, v0 = cases.cylinder(nx=NX, ny=NY, pressure_solver="mgcg", poisson_tol=TOL,
                                       poisson_maxiter=300, precision=precision)
    sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
    assert np.array_equal(sim.solver.solid, case.solid)
    res = sim.run(max_steps=STEPS)
    scale = np.abs(u_ref).max()
    du = np.abs(sim.fields.u - u_ref).max() / scale
    dv = np.abs(sim.fields.v - v_ref).max() / scale
    divs = [float(d) for d in res.history["max_div"]]
    print(f"{precision}: max|du|/|u| {du:.2e}, max|dv|/|u| {dv:.2e}, max|div u| per step "
          + ", ".join(f"{d:.2e}" for d in divs)
          + f", {sim.solver.pressure_solver.last_iterations} it/solve", flush=True)
    assert du <= 1e-12 and dv <= 1e-12
    assert max(divs) <= 3e-12
    # velocities equal to ~2e-15 bound the divergence gap by ~4 * 2e-15 / dx
    assert all(abs(d - r) <= 1e-12 for d, r in zip(divs, div_ref))
