"""Fractional-step solver: the reference's ``ProjectionSolver`` on the GPU.

Same constructor, attributes and methods as ``pycfd/core/solver.py:82-756``.
Setup (face ranges, body face masks, Poisson operator, multigrid hierarchy)
happens once on the host; ``initialize``, ``project`` and ``step`` run
entirely in HIP kernels on device-resident FP64 fields.  ``momentum_rhs``,
``divergence`` and ``project`` also accept NumPy arrays, so tests written for
the reference can call them unchanged.
"""

from __future__ import annotations

import logging
import os
from collections.abc import Sequence

import numpy as np

from ..config import AdvectionScheme, BCKind, SimulationConfig, SolverType
from .boundary import BoundaryManager
from .engine import DeviceContext
from .fields import FlowField, device_empty, device_zeros, to_device, to_host
from .mesh import StructuredMesh
from .pressure import assemble_poisson_matrix, make_pressure_solver

log = logging.getLogger(__name__)


def _face_solid_mask(solid: np.ndarray, lo: int, hi: int, axis: int,
                     periodic: bool) -> np.ndarray:
    """Faces in ``[lo, hi)`` touching a solid cell; face m sits between cells
    m-2 and m-1 (solver.py:62-79)."""
    m = np.arange(lo, hi)
    if periodic:
        n = solid.shape[axis]
        return np.take(solid, (m - 2) % n, axis=axis) | np.take(solid, (m - 1) % n, axis=axis)
    pad = [(0, 0), (0, 0)]
    pad[axis] = (1, 1)
    padded = np.pad(solid, pad, constant_values=False)
    return np.take(padded, m - 1, axis=axis) | np.take(padded, m, axis=axis)


class ProjectionSolver:
    """Chorin projection with SSP-RK3 on the staggered grid, on one MI355X."""

    def __init__(self, config: SimulationConfig, mesh: StructuredMesh | None = None,
                 obstacle: np.ndarray | None = None,
                 bodies: Sequence[np.ndarray] | None = None, *, soft_body: bool = False) -> None:
        """``soft_body``: the obstacle mask only fixes the Poisson operator's
        topology, and the body is forced by a soft mask through the
        differentiable step (pycfd_amd.adjoint) instead of the binary face
        forcing; ``step`` then does not force the body at all."""
        if config.solver_type is SolverType.SIMPLE:
            raise NotImplementedError(
                "solver_type='simple' is reserved for a future SIMPLE "
                "(pressure-correction) implementation; only 'projection' is "
                "available. Set solver_type='projection'.")
        self.config = config
        self.mesh = mesh if mesh is not None else StructuredMesh.from_config(config)
        self.metrics = self.mesh.metrics
        self.nu = config.nu
        self.boundaries = BoundaryManager(config.boundary_config, self.mesh)
        self.periodic_x = self.boundaries.periodic_x
        self.periodic_y = self.boundaries.periodic_y
        self._reject_stretched_periodic()
        nx, ny = self.mesh.shape
        if obstacle is None:
            self.solid = np.zeros((nx, ny), dtype=bool)
        else:
            self.solid = np.asarray(obstacle, dtype=bool)
            if self.solid.shape != (nx, ny):
                raise ValueError(
                    f"obstacle mask must have shape {(nx, ny)}, got {self.solid.shape}")
        self.has_obstacle = bool(self.solid.any())
        self.body_masks: tuple[np.ndarray, ...] = ()
        if bodies is not None and len(bodies) > 1:
            self.body_masks = tuple(np.asarray(b, dtype=bool) for b in bodies)
            for k, b in enumerate(self.body_masks):
                if b.shape != (nx, ny):
                    raise ValueError(f"body {k} mask must have shape {(nx, ny)}, got {b.shape}")
            self._check_bodies(nx, ny)
        self._build_slices()
        self.dirichlet_pressure = self.boundaries.dirichlet_pressure()
        self.poisson = assemble_poisson_matrix(
            self.mesh, self.periodic_x, self.periodic_y,
            self.solid if self.has_obstacle else None,
            dirichlet=self.dirichlet_pressure or None)
        self.turbulence = None
        if config.use_les:  # solver.py:164-167
            from ..physics.turbulence import SmagorinskyModel
            self.turbulence = SmagorinskyModel(self.mesh, config.smagorinsky_cs)
        self._use_kernel = self.turbulence is None
        self._blend = 0.0

        walls = ("left", "right", "bottom", "top")
        conds = self.boundaries.conditions
        body_id = None
        if self.body_masks:
            body_id = np.full((nx, ny), -1, dtype=np.int8)
            for k, b in enumerate(self.body_masks):
                body_id[b] = k
        stretched = not self.mesh.is_uniform
        self.ctx = DeviceContext(
            nx, ny, float(self.mesh.dx_cells[0]), float(self.mesh.dy_cells[0]), self.nu,
            config.body_force,
            config.time_scheme.value,
            [conds[w].kind.value for w in walls],
            [conds[w].spec.velocity for w in walls],
            [conds[w].spec.p_ref for w in walls],
            [conds[w].inlet_line(self.mesh) for w in walls],
            self.solid if self.has_obstacle else None, body_id, len(self.body_masks),
            les_coeff=self.turbulence.device_coeff if self.turbulence is not None else None,
            les_cs=self.turbulence.cs if self.turbulence is not None else 0.0,
            dx_cells=self.mesh.dx_cells if stretched else None,
            dy_cells=self.mesh.dy_cells if stretched else None,
            area_mean=self.mesh.area_mean if stretched else None, soft_ib=soft_body)
        if self.turbulence is not None:
            self.turbulence.bind(self.ctx, self.nu)
        self.pressure_solver = make_pressure_solver(
            config.pressure_solver, self.poisson, tol=config.poisson_tol,
            maxiter=config.poisson_maxiter, sor_omega=config.sor_omega,
            mg_sweeps=config.mg_sweeps, context=self.ctx, precision=config.precision,
            mg_coarse_max=config.mg_coarse_max)
        self.pressure_solver.configure()
        self._use_graphs = False
        self.use_graphs = os.environ.get("PCFD_STEP_GRAPHS", "0") == "1"
        self._warm_order = 0
        self._has_guess = False
        self._force = np.zeros(2 + 2 * len(self.body_masks))
        self._force_stale = False

    MAX_BODIES = 128  # the device keeps one signed byte of body index per cell

    def _check_bodies(self, nx: int, ny: int) -> None:
        """Each face is charged to exactly one body (solver.py:289-296 builds
        disjoint face sets for bodies that neither overlap nor touch), so
        bodies that overlap or share a face are refused, as are more bodies
        than the device's body index holds."""
        if len(self.body_masks) > self.MAX_BODIES:
            raise ValueError(f"at most {self.MAX_BODIES} bodies are supported, "
                             f"got {len(self.body_masks)}")
        ids = np.full((nx, ny), -1, dtype=np.int32)
        for k, b in enumerate(self.body_masks):
            if np.any(ids[b] >= 0):
                raise ValueError(f"body {k} overlaps body {int(ids[b].max())}")
            ids[b] = k
        for axis, periodic in ((0, self.periodic_x), (1, self.periodic_y)):
            a = ids
            b = np.roll(ids, -1, axis=axis)
            pair = (a >= 0) & (b >= 0) & (a != b)
            if not periodic:  # the wrap-around pair is not a face
                last = [slice(None)] * 2
                last[axis] = -1
                pair[tuple(last)] = False
            if pair.any():
                i = np.argwhere(pair)[0]
                raise ValueError(f"bodies {int(a[tuple(i)])} and {int(b[tuple(i)])} touch: "
                                 "they share a face with no fluid between them")

    # ------------------------------------------------------------------ #
    def _build_slices(self) -> None:
        """Solvable face ranges and body face masks (solver.py:235-296)."""
        nx, ny = self.mesh.shape
        kinds = {w: bc.kind for w, bc in self.boundaries.conditions.items()}
        outlet = BCKind.PRESSURE_OUTLET
        u_lo = 1 if (self.periodic_x or kinds["left"] is outlet) else 2
        u_hi = nx + 2 if (kinds["right"] is outlet and not self.periodic_x) else nx + 1
        self.u_upd = (slice(u_lo, u_hi), slice(1, ny + 1))
        self.u_rhs_sel = slice(u_lo - 1, u_hi - 1)
        v_lo = 1 if (self.periodic_y or kinds["bottom"] is outlet) else 2
        v_hi = ny + 2 if (kinds["top"] is outlet and not self.periodic_y) else ny + 1
        self.v_upd = (slice(1, nx + 1), slice(v_lo, v_hi))
        self.v_rhs_sel = slice(v_lo - 1, v_hi - 1)
        if self.has_obstacle:
            self.u_face_solid = _face_solid_mask(self.solid, u_lo, u_hi, 0, self.periodic_x)
            self.v_face_solid = _face_solid_mask(self.solid, v_lo, v_hi, 1, self.periodic_y)
        else:
            self.u_face_solid = None
            self.v_face_solid = None
        self.u_face_body = tuple(_face_solid_mask(b, u_lo, u_hi, 0, self.periodic_x)
                                 for b in self.body_masks)
        self.v_face_body = tuple(_face_solid_mask(b, v_lo, v_hi, 1, self.periodic_y)
                                 for b in self.body_masks)

    # -- captured steps ------------------------------------------------- #
    @property
    def use_graphs(self) -> bool:
        """Whether ``step`` replays captured HIP graphs (the same arithmetic
        as launching each kernel; see pcfd_set_step_graphs).  Off unless the
        environment sets ``PCFD_STEP_GRAPHS=1``: on MI355X a replay runs its
        kernels no faster than the stream does (the per-kernel dispatch on
        the device, not the host's launches, sets the pace)."""
        return self._use_graphs

    @use_graphs.setter
    def use_graphs(self, enable: bool) -> None:
        self.ctx.set_step_graphs(bool(enable))
        self._use_graphs = bool(enable)

    @property
    def warm_start_order(self) -> int:
        """How each substep's pressure solve is started: 0 (the default, the
        reference's ``_p_guess``) from the last pressure solved; 1-4 from the
        polynomial extrapolation in time of the same substep's last 2-5
        solutions, equal steps assumed (2 p1 - p2; 3 p1 - 3 p2 + p3;
        4 p1 - 6 p2 + 4 p3 - p4; 5 p1 - 10 p2 + 10 p3 - 5 p4 + p5).  Weights
        from the actual step times change nothing measurable (the adaptive
        dt varies by a fraction of a percent a step), and orders above 4
        saved no further iterations.  The solves still stop on the same
        tolerance; a closer start only saves iterations (on the 2048x1024
        cylinder 11 -> 6 per solve at order 4 once five steps are stored).
        Resetting or setting the pressure guess restarts the history."""
        return self._warm_order

    @warm_start_order.setter
    def warm_start_order(self, order: int) -> None:
        order = int(order)
        if not 0 <= order <= 4:
            raise ValueError(f"warm_start_order must be 0-4, got {order}")
        self.ctx.set_warm_start(order)
        self._warm_order = order

    # -- forces --------------------------------------------------------- #
    @property
    def body_force_reaction(self) -> tuple[float, float]:
        """Reaction force of the last substep (solver.py:173-174)."""
        f = self._current_force()
        return float(f[0]), float(f[1])

    @property
    def body_force_reactions(self) -> tuple[tuple[float, float], ...]:
        f = self._current_force()
        return tuple((float(f[2 + 2 * k]), float(f[3 + 2 * k]))
                     for k in range(len(self.body_masks)))

    def _current_force(self) -> np.ndarray:
        if self._force_stale:
            self._force = self.ctx.force()
            self._force_stale = False
        return self._force

    def set_force(self, fx: float, fy: float) -> None:
        """Total force read back together with the step diagnostics."""
        if self.body_masks:
            self._force_stale = True
        else:
            self._force[0], self._force[1] = fx, fy
            self._force_stale = False

    # -- initialisation --------------------------------------------------- #
    def initialize(self, u_init=None, v_init=None) -> FlowField:
        """Initial state with BCs applied and projected (solver.py:301-329)."""
        nx, ny = self.mesh.shape
        u = np.zeros(self.mesh.u_shape)
        v = np.zeros(self.mesh.v_shape)
        if u_init is not None:
            u[1:nx + 2, 1:ny + 1] = u_init
        if v_init is not None:
            v[1:nx + 1, 1:ny + 2] = v_init
        fields = FlowField(self.mesh, to_device(u), to_device(v), device_zeros(self.mesh.p_shape))
        info = self.ctx.initialize(fields.u_dev, fields.v_dev, fields.p_dev)
        self.pressure_solver.record(info.iterations, info.residual, bool(info.converged))
        self._has_guess = True
        return fields

    # -- spatial operators ------------------------------------------------ #
    def momentum_rhs(self, u, v, blend: float = 0.0):
        """``-A(u) + nu lap(u) + f`` on the GPU (solver.py:501-529)."""
        import torch
        nx, ny = self.mesh.shape
        host = not isinstance(u, torch.Tensor)
        ud, vd = to_device(u), to_device(v)
        ru = torch.empty((nx + 1, ny), dtype=torch.float64, device="cuda")
        rv = torch.empty((nx, ny + 1), dtype=torch.float64, device="cuda")
        self.ctx.momentum_rhs(ud, vd, blend, ru, rv)
        return (to_host(ru), to_host(rv)) if host else (ru, rv)

    def momentum_rhs_variable(self, u, v, nu_c, blend: float = 0.0):
        """The RHS with the stress-divergence viscous term of a given
        cell-centred viscosity (solver.py:467-499 inside momentum_rhs)."""
        import torch
        nx, ny = self.mesh.shape
        host = not isinstance(u, torch.Tensor)
        ru = torch.empty((nx + 1, ny), dtype=torch.float64, device="cuda")
        rv = torch.empty((nx, ny + 1), dtype=torch.float64, device="cuda")
        self.ctx.momentum_rhs_variable(to_device(u), to_device(v), blend, to_device(nu_c), ru, rv)
        return (to_host(ru), to_host(rv)) if host else (ru, rv)

    def divergence(self, u, v):
        """Cell-centred divergence ``(nx, ny)`` (solver.py:532-543)."""
        import torch
        nx, ny = self.mesh.shape
        host = not isinstance(u, torch.Tensor)
        out = torch.empty((nx, ny), dtype=torch.float64, device="cuda")
        self.ctx.divergence(to_device(u), to_device(v), out)
        return to_host(out) if host else out

    def max_divergence(self, fields: FlowField) -> float:
        """Largest |div| over the fluid cells (solver.py:545-550)."""
        d = fields.diagnostics
        if "max_div" in d:
            return d["max_div"]
        out = self.ctx.diagnostics(fields.u_dev, fields.v_dev, None)
        return float(out[5])

    def project(self, u, v, dt: float):
        """Remove the divergence in place; return the ghosted pressure (solver.py:643-683)."""
        import torch
        host = not isinstance(u, torch.Tensor)
        ud, vd = to_device(u), to_device(v)
        p = device_zeros(self.mesh.p_shape)
        info = self.ctx.project(ud, vd, dt, p)
        self.pressure_solver.record(info.iterations, info.residual, bool(info.converged))
        self._has_guess = True
        if host:
            u[...] = to_host(ud)
            v[...] = to_host(vd)
            return to_host(p)
        if ud.data_ptr() != u.data_ptr():
            u.copy_(ud)
            v.copy_(vd)
        return p

    # -- time step -------------------------------------------------------- #
    def step(self, fields: FlowField, dt: float) -> FlowField:
        """Advance by ``dt`` with the configured scheme (solver.py:686-731)."""
        mesh = self.mesh
        # Every entry is written by the step: the last substep's predictor
        # covers the full ghosted u and v, the scatter the full ghosted p.
        u, v, p = device_empty(mesh.u_shape), device_empty(mesh.v_shape), device_empty(mesh.p_shape)
        info = self.ctx.step(fields.u_dev, fields.v_dev, dt, self._blend, u, v, p)
        self.pressure_solver.record(info.iterations, info.residual, bool(info.converged))
        self._has_guess = True
        self._force_stale = self.has_obstacle
        return FlowField(mesh, u, v, p, fields.t + dt, fields.step + 1)

    # -- warm start (solver.py:171, 662-663) ------------------------------ #
    def pressure_guess(self) -> np.ndarray | None:
        """The normalised pressure the next solve starts from, ``(nx, ny)``,
        or ``None`` before the first projection."""
        if not self._has_guess:
            return None
        import torch
        buf = torch.empty(self.mesh.shape, dtype=torch.float64, device="cuda")
        self.ctx.pressure_guess(buf, store=False)
        return to_host(buf)

    def set_pressure_guess(self, p: np.ndarray | None) -> None:
        """Replace the warm-start pressure (``None`` forgets it)."""
        if p is None:
            self.ctx.reset_guess()
            self._has_guess = False
            return
        arr = np.asarray(p, dtype=np.float64)
        if arr.shape != tuple(self.mesh.shape):
            raise ValueError(f"pressure guess has shape {arr.shape}, expected {self.mesh.shape}")
        self.ctx.pressure_guess(to_device(arr), store=True)
        self._has_guess = True

    # -- in-project tuning ------------------------------------------------ #
    def tune_smoother_sweeps(self, fields: FlowField, dt: float, candidates=(1, 2, 3),
                             steps: int = 3, margin: float = 0.03) -> dict[int, dict]:
        """Pick the multigrid smoothing sweeps (V(s, s)) that make a step
        fastest, by measurement: for each candidate the engine is configured
        with it and `steps` whole steps are timed from the same `fields` and
        the same warm-start pressure; the fewest sweeps within `margin` of
        the fastest stay configured (a few steps' timings differ by a percent
        or two from run to run, and fewer sweeps is the cheaper cycle when
        they tie).  The reference exposes the choice as ``mg_sweeps`` and documents the
        trade-off (more sweeps, fewer iterations, docs/usage.md:473-476);
        this measures it on the device for the case at hand.  Only the
        multigrid solvers have sweeps; returns ``{sweeps: {"ms_per_step",
        "iterations"}}``, empty for the others.  Setting the pressure guess
        before each timed step restarts the warm-start history, so every
        candidate is timed with solves started from the last pressure."""
        import time

        import torch
        from .pressure import MultigridSolverBase
        ps = self.pressure_solver
        if not isinstance(ps, MultigridSolverBase):
            return {}
        guess = self.pressure_guess()
        graphs = self.use_graphs
        self.use_graphs = False
        table: dict[int, dict] = {}
        for sw in candidates:
            ps.smooth_sweeps = int(sw)
            ps.configure()
            times, its = [], 0
            for rep in range(steps + 1):  # the first step warms the tuned plan up
                self.set_pressure_guess(guess)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                self.step(fields, dt)
                torch.cuda.synchronize()
                if rep:
                    times.append(time.perf_counter() - t0)
                its = ps.last_iterations
            table[int(sw)] = {"ms_per_step": 1e3 * min(times), "iterations": int(its)}
        fastest = min(v["ms_per_step"] for v in table.values())
        best = min(k for k, v in table.items() if v["ms_per_step"] <= fastest * (1.0 + margin))
        ps.smooth_sweeps = best
        ps.configure()
        self.set_pressure_guess(guess)
        self.use_graphs = graphs
        log.info("smoother sweeps tuned: %s -> V(%d,%d)", table, best, best)
        return table

    # ------------------------------------------------------------------ #
    def set_upwind_blend(self, blend: float) -> None:
        self._blend = float(np.clip(blend, 0.0, 1.0))

    def upwind_blend_for(self, fields: FlowField, dt: float) -> float:
        """solver.py:738-756 (the worst cell: smallest widths)."""
        if self.config.advection_scheme is AdvectionScheme.CENTRAL:
            return 0.0
        if self.config.upwind_blend is not None:
            return self.config.upwind_blend
        umax, vmax = fields.max_velocity()
        m = self.metrics
        courant = max(umax * dt / m.min_hx, vmax * dt / m.min_hy)
        return float(min(1.0, 1.2 * courant))

    def _reject_stretched_periodic(self) -> None:
        """solver.py:209-230: a periodic axis must keep equal first and last cells."""
        mesh = self.mesh
        for axis, stretched, periodic, ratio in (
                ("x", mesh.stretched_x, self.periodic_x, mesh.stretch_x),
                ("y", mesh.stretched_y, self.periodic_y, mesh.stretch_y)):
            if stretched and periodic:
                raise ValueError(
                    f"the {axis} axis is periodic and stretched (stretch_{axis}={ratio}): "
                    "the first and last cells differ in width, so the wrap would join "
                    f"mismatched cells. Use stretch_{axis}=1.0 on a periodic axis, or a "
                    "non-periodic boundary on that axis.")


__all__ = ["ProjectionSolver", "_face_solid_mask"]
