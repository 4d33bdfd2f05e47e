"""Time loop: adaptive dt, health checks, per-step diagnostics.

Mirrors ``pycfd/core/timestepper.py:47-266``.  Every full-field reduction the
reference makes per step (max|u|, max|v|, finiteness, max|du|, max|div|,
kinetic energy) comes from ONE fused device pass (``pcfd_diagnostics``) read
back together with the body force, so the host touches a handful of scalars
per step.
"""

from __future__ import annotations

import logging
import time
from dataclasses import dataclass, field
from typing import Callable

from ..config import (DIVERGENCE_VELOCITY_LIMIT, MIN_TIME_STEP, QUIESCENT_VELOCITY_FLOOR,
                      VISCOUS_SAFETY_FACTOR, SimulationConfig)
from .fields import FlowField
from .solver import ProjectionSolver

log = logging.getLogger(__name__)


class DivergenceError(RuntimeError):
    """The solution blew up or produced non-finite values (timestepper.py:47-48)."""


@dataclass
class SimulationResult:
    fields: FlowField
    steps: int
    time: float
    wall_time: float
    converged: bool = False
    history: dict[str, list] = field(default_factory=dict)

    def summary(self) -> str:
        rate = self.steps / self.wall_time if self.wall_time > 0 else float("nan")
        status = "steady state reached" if self.converged else "end time reached"
        return (f"{status}: {self.steps} steps, t = {self.time:.4f}, "
                f"{self.wall_time:.2f} s wall ({rate:.0f} steps/s)")


class TimeStepper:
    """Runs a :class:`ProjectionSolver` forward in time (timestepper.py:72-266)."""

    def __init__(self, solver: ProjectionSolver, config: SimulationConfig | None = None) -> None:
        self.solver = solver
        self.config = config if config is not None else solver.config
        self.mesh = solver.mesh
        self.viscous_dt_limit = self._viscous_limit()

    def _viscous_limit(self) -> float:
        nu = self.solver.nu
        if nu <= 0:
            return float("inf")
        m = self.mesh.metrics
        dx, dy = m.min_hx, m.min_hy
        return VISCOUS_SAFETY_FACTOR / (2.0 * nu * (1.0 / dx ** 2 + 1.0 / dy ** 2))

    def measure(self, fields: FlowField, previous: FlowField | None = None) -> dict:
        """Fused device reductions of ``fields`` (and its change since ``previous``)."""
        out = self.solver.ctx.diagnostics(
            fields.u_dev, fields.v_dev, fields.p_dev,
            previous.u_dev if previous is not None else None,
            previous.v_dev if previous is not None else None)
        d = fields.diagnostics
        d["umax"], d["vmax"] = float(out[0]), float(out[1])
        d["nonfinite"] = int(out[2])
        d["du"] = max(float(out[3]), float(out[4]))
        d["max_div"] = float(out[5])
        # mesh.py:213-223: one area after the sum on a uniform mesh; on a
        # stretched one the device already weighted each cell by its area
        area = self.mesh.metrics.uniform_area
        d["kinetic_energy"] = 0.5 * (float(out[6]) * area if area is not None
                                     else float(out[6]))
        self.solver.set_force(float(out[7]), float(out[8]))
        return d

    def compute_dt(self, fields: FlowField) -> float:
        """timestepper.py:94-113."""
        if not self.config.adaptive_dt:
            return self.config.dt
        umax, vmax = fields.max_velocity()
        m = self.mesh.metrics
        scale = umax / m.min_hx + vmax / m.min_hy
        dt_conv = self.config.cfl_max / scale if scale > QUIESCENT_VELOCITY_FLOOR else float("inf")
        return float(min(dt_conv, self.viscous_dt_limit, self.config.dt))

    def cfl_number(self, fields: FlowField, dt: float) -> float:
        umax, vmax = fields.max_velocity()
        m = self.mesh.metrics
        return float(dt * (umax / m.min_hx + vmax / m.min_hy))

    def _check_health(self, fields: FlowField, dt: float) -> None:
        """timestepper.py:122-142."""
        if not fields.is_finite():
            raise DivergenceError(
                f"non-finite values at step {fields.step} (t = {fields.t:.6g}). "
                "The most common causes are a time step above the stability "
                "limit and a Reynolds number too high for the grid; try "
                "adaptive_dt=True, a smaller cfl_max, or advection_scheme='upwind'.")
        umax, vmax = fields.max_velocity()
        if max(umax, vmax) > DIVERGENCE_VELOCITY_LIMIT:
            raise DivergenceError(
                f"solution diverged at step {fields.step} (t = {fields.t:.6g}): "
                f"max|u| = {umax:.3e}, max|v| = {vmax:.3e} exceeds the limit "
                f"{DIVERGENCE_VELOCITY_LIMIT:.0e}.")
        if dt < MIN_TIME_STEP:
            raise DivergenceError(
                f"adaptive time step collapsed to {dt:.3e} (< {MIN_TIME_STEP:.0e}) "
                f"at step {fields.step}; the solution is probably diverging.")

    def run(self, fields: FlowField, t_end: float | None = None, max_steps: int | None = None,
            callback: Callable[[FlowField, dict], None] | None = None,
            callback_every: int = 50, progress: bool = False,
            record_every: int = 1) -> SimulationResult:
        """Integrate until ``t_end`` (timestepper.py:145-248)."""
        t_end = self.config.t_end if t_end is None else t_end
        max_steps = self.config.max_steps if max_steps is None else max_steps
        history: dict[str, list] = {k: [] for k in ("step", "t", "dt", "cfl", "max_div",
                                                    "kinetic_energy", "residual", "fx", "fy")}
        if "umax" not in fields.diagnostics:
            self.measure(fields)
        wall_start = time.perf_counter()
        converged = False
        steps_taken = 0
        while fields.t < t_end - 1e-12:
            if max_steps is not None and steps_taken >= max_steps:
                break
            dt = self.compute_dt(fields)
            self._check_health(fields, dt)
            dt = min(dt, t_end - fields.t)
            self.solver.set_upwind_blend(self.solver.upwind_blend_for(fields, dt))
            previous = fields
            fields = self.solver.step(previous, dt)
            steps_taken += 1
            d = self.measure(fields, previous)
            rate_of_change = d["du"] / dt
            if steps_taken % record_every == 0:
                info = self._diagnostics(fields, dt, rate_of_change)
                for k, val in info.items():
                    if k in history:
                        history[k].append(val)
                fields.diagnostics.update(info)
            if callback is not None and steps_taken % callback_every == 0:
                callback(fields, self._diagnostics(fields, dt, rate_of_change))
            if self.config.steady_tol is not None and rate_of_change < self.config.steady_tol:
                converged = True
                log.info("steady state at step %d (t = %.4f)", fields.step, fields.t)
                break
        self._check_health(fields, self.config.dt)
        wall = time.perf_counter() - wall_start
        if callback is not None:
            callback(fields, self._diagnostics(fields, self.config.dt, float("nan")))
        result = SimulationResult(fields, steps_taken, fields.t, wall, converged, history)
        log.info(result.summary())
        return result

    def _diagnostics(self, fields: FlowField, dt: float, rate_of_change: float) -> dict:
        """timestepper.py:251-266."""
        fx, fy = self.solver.body_force_reaction
        d = fields.diagnostics
        return {
            "step": fields.step, "t": fields.t, "dt": dt,
            "cfl": self.cfl_number(fields, dt),
            "max_div": d["max_div"],
            "kinetic_energy": d["kinetic_energy"],
            "residual": self.solver.pressure_solver.last_residual,
            "rate_of_change": rate_of_change,
            "fx": fx, "fy": fy,
        }
