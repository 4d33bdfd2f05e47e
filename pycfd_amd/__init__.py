"""pycfd_amd -- pycfd's projection time step on the AMD Instinct MI355X.

The staggered-grid projection step of pycfd (SSP-RK3 advection/diffusion,
immersed-boundary forcing, divergence -> pressure Poisson -> gradient
correction, and the CG / red-black SOR / aggregation-multigrid solvers) runs
as hand-written gfx950 HIP kernels in ``libpycfd_hip.so``; this package is the
host side that keeps the reference's API:

    from pycfd_amd import Simulation, SimulationConfig
    sim = Simulation(SimulationConfig(pressure_solver="mgcg", ...))
    result = sim.run(max_steps=100)
"""

__version__ = "0.1.0"

from .config import (AdvectionScheme, BCKind, BCSpec, PressureSolver, SimulationConfig,
                     SolverType, TimeScheme)

__all__ = [
    "AdvectionScheme", "BCKind", "BCSpec", "PressureSolver", "SimulationConfig", "SolverType",
    "TimeScheme", "Simulation", "ProjectionSolver",
]


def __getattr__(name):
    # The GPU-facing classes import torch; keep `import pycfd_amd` light.
    if name == "Simulation":
        from .physics.incompressible import Simulation
        return Simulation
    if name == "ProjectionSolver":
        from .core.solver import ProjectionSolver
        return ProjectionSolver
    raise AttributeError(name)
