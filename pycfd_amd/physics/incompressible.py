"""``Simulation``: the drop-in driver (pycfd/physics/incompressible.py:29-256).

``Simulation(config, obstacle, u_init, v_init).run(...)`` builds the mesh,
the GPU ``ProjectionSolver`` and the ``TimeStepper`` exactly as the reference
wires them, and exposes the same diagnostics and force coefficients.
"""

from __future__ import annotations

import logging
from collections.abc import Sequence
from typing import Callable

import numpy as np

from ..analysis.postprocess import Probe, force_coefficients, kinetic_energy, vorticity
from ..config import SimulationConfig
from ..core.fields import FlowField
from ..core.mesh import StructuredMesh
from ..core.solver import ProjectionSolver
from ..core.timestepper import SimulationResult, TimeStepper
from ..geometry.obstacles import Obstacle, ObstacleGroup, as_obstacle_group

log = logging.getLogger(__name__)


class Simulation:
    """A configured, ready-to-run incompressible flow problem on one MI355X."""

    def __init__(self, config: SimulationConfig,
                 obstacle: Obstacle | ObstacleGroup | Sequence[Obstacle] | None = None,
                 u_init=None, v_init=None) -> None:
        self.config = config
        self.mesh = StructuredMesh.from_config(config)
        self.bodies = as_obstacle_group(obstacle)
        self.obstacle = obstacle if isinstance(obstacle, (Obstacle, ObstacleGroup)) \
            else self.bodies
        self.solver = ProjectionSolver(
            config, self.mesh,
            self.bodies.mask if self.bodies is not None else None,
            bodies=self.bodies.masks if self.bodies is not None else None)
        self.stepper = TimeStepper(self.solver, config)
        self.fields = self.solver.initialize(u_init, v_init)
        self.probes: list[Probe] = []
        self.result: SimulationResult | None = None
        log.info("simulation ready\n%s", config.summary())

    @classmethod
    def from_checkpoint(cls, path,
                        obstacle: Obstacle | ObstacleGroup | Sequence[Obstacle] | None = None
                        ) -> "Simulation":
        """Rebuild a simulation from a checkpoint and restore its state
        (incompressible.py:74-84); a stored warm-start pressure is restored too."""
        from ..analysis.export import load_checkpoint_with_guess
        fields, config, guess = load_checkpoint_with_guess(path)
        sim = cls(config, obstacle)
        sim.fields = fields
        if guess is not None:
            sim.solver.set_pressure_guess(guess)
        log.info("resumed from %s at t=%.6g (step %d)", path, fields.t, fields.step)
        return sim

    def save_checkpoint(self, path):
        """Checkpoint the current state, the configuration and the warm-start
        pressure (what the reference's CLI writes with ``--checkpoint``)."""
        from ..analysis.export import save_checkpoint
        return save_checkpoint(self.fields, self.config, path, solver=self.solver)

    def add_probe(self, x: float, y: float, name: str = "probe") -> Probe:
        probe = Probe(x, y, name)
        self.probes.append(probe)
        return probe

    def run(self, t_end: float | None = None, max_steps: int | None = None,
            callback: Callable[[FlowField, dict], None] | None = None,
            callback_every: int = 50, progress: bool = False) -> SimulationResult:
        """Integrate forward, recording probes as it goes (incompressible.py:94-116)."""

        def _callback(fields: FlowField, info: dict) -> None:
            for probe in self.probes:
                probe.record(fields)
            if callback is not None:
                callback(fields, info)

        cb = _callback if (self.probes or callback is not None) else None
        self.result = self.stepper.run(self.fields, t_end=t_end, max_steps=max_steps,
                                       callback=cb, callback_every=callback_every,
                                       progress=progress)
        self.fields = self.result.fields
        return self.result

    @property
    def solid_mask(self) -> np.ndarray | None:
        return self.bodies.mask if self.bodies is not None else None

    def force_coefficients(self) -> tuple[float, float]:
        """(Cd, Cl) from the immersed-boundary reaction (incompressible.py:126-142)."""
        if self.bodies is None:
            return 0.0, 0.0
        return force_coefficients(self.solver.body_force_reaction, self.config.u_ref,
                                  self.config.l_ref)

    def force_coefficients_by_body(self) -> dict[str, tuple[float, float]]:
        if self.bodies is None:
            return {}
        reactions = self.solver.body_force_reactions or (self.solver.body_force_reaction,)
        return {name: force_coefficients(f, self.config.u_ref, self.config.l_ref)
                for name, f in zip(self.bodies.names, reactions)}

    def outlet_pressure_deviation(self) -> float:
        walls = self.solver.dirichlet_pressure
        if not walls:
            return float("nan")
        from ..core.boundary import wall_index
        worst = 0.0
        p = self.fields.p
        for wall, p_ref in walls.items():
            wi = wall_index(wall, self.mesh)
            ghost = p[wi.p_ghost, :] if wi.axis == 0 else p[:, wi.p_ghost]
            inner = p[wi.p_in, :] if wi.axis == 0 else p[:, wi.p_in]
            worst = max(worst, float(np.abs(0.5 * (ghost + inner) - p_ref).max()))
        return worst

    def diagnostics(self) -> dict[str, float]:
        """incompressible.py:193-208."""
        mask = self.solid_mask
        cd, cl = self.force_coefficients()
        out = {
            "t": self.fields.t,
            "step": self.fields.step,
            "kinetic_energy": kinetic_energy(self.fields, mask),
            "max_divergence": self.solver.max_divergence(self.fields),
            "max_vorticity": float(np.abs(vorticity(self.fields)).max()),
            "cd": cd,
            "cl": cl,
        }
        if self.solver.dirichlet_pressure:
            out["outlet_pressure_dev"] = self.outlet_pressure_deviation()
        return out

    def __repr__(self) -> str:
        return (f"Simulation({self.config.name!r}, {self.mesh.nx}x{self.mesh.ny}, "
                f"Re={self.config.re:g}, t={self.fields.t:.4g})")
