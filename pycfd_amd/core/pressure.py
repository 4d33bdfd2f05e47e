"""Pressure Poisson system and the solver-selection API.

``assemble_poisson_matrix`` builds the reference's operator on the host
(pycfd/core/pressure.py:110-270) for the multigrid setup and for
introspection; the device applies the same operator matrix-free from one
flag byte per cell (csrc/poisson.hip).  ``make_pressure_solver`` and the
solver classes keep the reference's interface (pressure.py:276-660):
``solve(rhs, p0=None) -> (nx, ny)``, ``last_iterations``, ``last_residual``,
``warm_startable``, ``system``, ``tol``, ``maxiter``, ``name`` -- every solve
runs on the GPU.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from ..config import (DEFAULT_MG_COARSE_MAX, DEFAULT_MG_SWEEPS, DEFAULT_POISSON_MAXITER,
                      DEFAULT_POISSON_TOL, DEFAULT_SOR_OMEGA, PressureSolver)
from .mesh import StructuredMesh

_ZERO_RHS_NORM = 1.0e-300

#: Damping of the standalone Jacobi solver (pressure.py:407).
JACOBI_DAMPING = 2.0 / 3.0

#: Tolerance the ``direct`` selection is solved to.  The GPU path has no sparse
#: LU; it runs multigrid-preconditioned CG to this relative residual, which is
#: at the level of a direct solve's round-off (pressure.py:330-363 reaches
#: ~1e-14).
DIRECT_TOL = 1.0e-13
DIRECT_MAXITER = 200


@dataclass
class PoissonSystem:
    """Assembled operator plus what the device needs to rebuild it."""

    A: sp.csr_matrix
    shape: tuple[int, int]
    fluid: np.ndarray
    singular: bool
    pin: int
    bc_rhs: np.ndarray = None  # type: ignore[assignment]
    row_scale: float | np.ndarray = 1.0
    dx: float = 1.0
    dy: float = 1.0
    periodic_x: bool = False
    periodic_y: bool = False
    solid: np.ndarray | None = None
    dirichlet: dict = field(default_factory=dict)
    # stretched mesh: cell widths and mean cell area (None on a uniform mesh)
    dx_cells: np.ndarray | None = None
    dy_cells: np.ndarray | None = None
    area_mean: float | None = None

    def __post_init__(self) -> None:
        if self.bc_rhs is None:
            self.bc_rhs = np.zeros(self.A.shape[0])

    @property
    def n(self) -> int:
        return self.A.shape[0]


def assemble_poisson_matrix(mesh: StructuredMesh, periodic_x: bool = False,
                            periodic_y: bool = False, solid_mask: np.ndarray | None = None,
                            dirichlet: dict[str, float] | None = None) -> PoissonSystem:
    """5-point ``div(grad)`` over the correctable faces (pressure.py:110-270).

    On a stretched mesh each face carries face area / centre distance / mean
    cell area and each row is scaled by its cell area / mean cell area
    (pressure.py:193-207), which keeps the operator symmetric.
    """
    nx, ny = mesh.shape
    n = nx * ny
    solid = (np.zeros((nx, ny), dtype=bool) if solid_mask is None
             else np.asarray(solid_mask, dtype=bool))
    if solid.shape != (nx, ny):
        raise ValueError(f"solid_mask must have shape {(nx, ny)}, got {solid.shape}")
    fluid = ~solid
    dirichlet = dict(dirichlet or {})
    stretched = not mesh.is_uniform
    if stretched:
        m = mesh.metrics
        vbar = mesh.area_mean
        cxf = m.hy / (m.hxu * vbar)   # (nx+1, ny), by x face
        cyf = m.hx / (m.hyv * vbar)   # (nx, ny+1), by y face
        row_scale = (m.cell_areas / vbar).ravel()
        dx, dy = float(mesh.dx_cells[0]), float(mesh.dy_cells[0])
    else:
        dx, dy = mesh.dx, mesh.dy
        cx, cy = 1.0 / (dx * dx), 1.0 / (dy * dy)
        row_scale = 1.0
    # the rows are assembled natively (csrc/assembly.hip)
    import ctypes as C

    from .. import _lib
    desc = _lib.PoissonAssembly()
    desc.nx, desc.ny = nx, ny
    desc.periodic_x, desc.periodic_y = int(bool(periodic_x)), int(bool(periodic_y))
    keep = []
    if solid.any():
        sarr = np.ascontiguousarray(solid, dtype=np.uint8)
        keep.append(sarr)
        desc.solid = sarr.ctypes.data_as(C.POINTER(C.c_uint8))
    if stretched:
        fx_ = np.ascontiguousarray(cxf, dtype=np.float64)
        fy_ = np.ascontiguousarray(cyf, dtype=np.float64)
        keep += [fx_, fy_]
        desc.cx_face = fx_.ctypes.data_as(C.POINTER(C.c_double))
        desc.cy_face = fy_.ctypes.data_as(C.POINTER(C.c_double))
    else:
        desc.cx, desc.cy = cx, cy
    for w, wall in enumerate(("left", "right", "bottom", "top")):
        desc.dirichlet[w] = int(wall in dirichlet)
        desc.p_ref[w] = float(dirichlet.get(wall, 0.0))
    lib = _lib.load()
    indptr = np.empty(n + 1, dtype=np.int64)
    _lib.check(lib.pcfd_poisson_assemble(C.byref(desc), indptr.ctypes.data, None, None, None,
                                         None), "pcfd_poisson_assemble")
    nnz = int(indptr[n])
    indices = np.empty(nnz, dtype=np.int32)
    data = np.empty(nnz)
    bc = np.empty(n)
    solvable = np.empty(n, dtype=np.uint8)
    _lib.check(lib.pcfd_poisson_assemble(C.byref(desc), indptr.ctypes.data, indices.ctypes.data,
                                         data.ctypes.data, bc.ctypes.data,
                                         solvable.ctypes.data), "pcfd_poisson_assemble")
    A = sp.csr_matrix((data, indices, indptr.astype(np.int32) if nnz < 2 ** 31 else indptr),
                      shape=(n, n))
    solvable = solvable.astype(bool).reshape(nx, ny)
    flat = solvable.ravel()
    pin = int(np.flatnonzero(flat)[0]) if flat.any() else 0
    return PoissonSystem(A, (nx, ny), flat.copy(), not dirichlet, pin, bc.ravel().copy(),
                         row_scale, dx, dy, bool(periodic_x), bool(periodic_y),
                         solid if solid.any() else None, dirichlet,
                         mesh.dx_cells.copy() if stretched else None,
                         mesh.dy_cells.copy() if stretched else None,
                         vbar if stretched else None)


def _standalone_context(system: PoissonSystem):
    """Device context for a solver used without a ProjectionSolver."""
    from .engine import DeviceContext
    kinds = []
    prefs = []
    for wall, periodic in (("left", system.periodic_x), ("right", system.periodic_x),
                           ("bottom", system.periodic_y), ("top", system.periodic_y)):
        if periodic:
            kinds.append("periodic")
        elif wall in system.dirichlet:
            kinds.append("pressure_outlet")
        else:
            kinds.append("no_slip")
        prefs.append(system.dirichlet.get(wall, 0.0))
    nx, ny = system.shape
    return DeviceContext(nx, ny, system.dx, system.dy, 1.0, bc_kinds=kinds, bc_p_ref=prefs,
                         solid=system.solid, dx_cells=system.dx_cells, dy_cells=system.dy_cells,
                         area_mean=system.area_mean)


class PoissonSolverBase:
    """Common interface (pressure.py:276-327); ``solve`` runs on the GPU."""

    name = "base"
    warm_startable = False
    _device_kind = "cg"

    def __init__(self, system: PoissonSystem, tol: float = DEFAULT_POISSON_TOL,
                 maxiter: int = DEFAULT_POISSON_MAXITER, *, context=None,
                 precision: str = "fp64", coarse_max: int = DEFAULT_MG_COARSE_MAX,
                 fused_levels: bool = True, tile: tuple[int, int] = (0, 0)) -> None:
        self.system = system
        self.tol = tol
        self.maxiter = maxiter
        self.precision = precision
        self.coarse_max = int(coarse_max)
        self.fused_levels = bool(fused_levels)
        self.tile = (int(tile[0]), int(tile[1]))
        # cooperative launch of the deepest multigrid levels: -1 never (the
        # default: on MI355X a grid-wide barrier costs more than the launch it
        # replaces), 0 measured choice, k >= 1 from level k down
        self.coarse_chain = -1
        # level 0's Krylov-start legs in streaming form: 0 measured choice,
        # -1 never, k > 0 always with k rows per wavefront
        self.stream_rows = 0
        # stored-level passes: 0 register-resident rows where they fit, -1 the
        # looping row kernels
        self.row_passes = 0
        # hand the stored multigrid levels over with 32-bit class indices even
        # when 16 bits would do (they are chosen by size otherwise)
        self.wide_classes = False
        self.last_iterations = 0
        self.last_residual = 0.0
        self._ctx = context
        self._configured = False

    # -- host helpers of the reference interface -----------------------
    def _prepare_rhs(self, rhs: np.ndarray) -> np.ndarray:
        b = (np.asarray(rhs, dtype=float).ravel() * self.system.row_scale
             + self.system.bc_rhs)
        fluid = self.system.fluid
        b[~fluid] = 0.0
        if self.system.singular:
            b[fluid] -= b[fluid].mean()
        return b

    def residual_norm(self, p: np.ndarray, b: np.ndarray) -> float:
        r = b - self.system.A @ p
        nb = float(np.linalg.norm(b))
        return float(np.linalg.norm(r)) / (nb if nb > _ZERO_RHS_NORM else 1.0)

    # -- device configuration ------------------------------------------
    def _levels(self):
        return None

    def _omega(self) -> float:
        return 1.0

    def _sweeps(self) -> int:
        return 1

    def _tol_maxiter(self) -> tuple[float, int]:
        return self.tol, self.maxiter

    @property
    def context(self):
        if self._ctx is None:
            self._ctx = _standalone_context(self.system)
        return self._ctx

    def configure(self) -> None:
        tol, maxiter = self._tol_maxiter()
        self.context.set_solver(self._device_kind, tol, maxiter, self._omega(), self._sweeps(),
                                1 if self.precision == "mixed" else 0, self._levels(),
                                fused_levels=self.fused_levels, tile=self.tile,
                                coarse_chain=self.coarse_chain, stream_rows=self.stream_rows,
                                row_passes=self.row_passes)
        self._configured = True

    def record(self, iterations: int, residual: float, converged: bool) -> None:
        """Store a device solve's outcome and raise like the reference."""
        self.last_iterations = int(iterations)
        self.last_residual = float(residual)
        self._check(converged)

    def _check(self, converged: bool) -> None:
        raise NotImplementedError

    def solve(self, rhs, p0=None) -> np.ndarray:
        """Solve ``lap(p) = rhs`` on the GPU; return ``p`` with shape ``(nx, ny)``."""
        from .fields import to_device, to_host
        import torch
        if not self._configured:
            self.configure()
        nx, ny = self.system.shape
        r = to_device(np.asarray(rhs, dtype=float).reshape(nx, ny) if not
                      isinstance(rhs, torch.Tensor) else rhs.reshape(nx, ny))
        x0 = None
        if p0 is not None:
            x0 = to_device(p0 if isinstance(p0, torch.Tensor) else
                           np.asarray(p0, dtype=float).reshape(nx, ny))
        out = torch.empty((nx, ny), dtype=torch.float64, device="cuda")
        info = self.context.pressure_solve(r, x0, out)
        self.record(info.iterations, info.residual, bool(info.converged))
        return to_host(out)


class _CGFamily(PoissonSolverBase):
    warm_startable = True
    _label = "pressure CG"

    def _check(self, converged: bool) -> None:
        if not converged and self.last_residual > max(self.tol * 10, 1e-8):
            raise RuntimeError(
                f"{self._label} failed to converge (info={self.maxiter}, "
                f"relative residual={self.last_residual:.3e} after "
                f"{self.last_iterations} iterations)"
            )


class CGSolver(_CGFamily):
    """Jacobi-preconditioned CG (pressure.py:366-398)."""

    name = "cg"
    _device_kind = "cg"


class JacobiSolver(PoissonSolverBase):
    """Damped Jacobi sweeps (pressure.py:410-454)."""

    name = "jacobi"
    warm_startable = True
    _device_kind = "jacobi"

    def __init__(self, system, tol=DEFAULT_POISSON_TOL, maxiter=DEFAULT_POISSON_MAXITER,
                 omega: float = JACOBI_DAMPING, **kw):
        super().__init__(system, tol, maxiter, **kw)
        self.omega = omega

    def _omega(self) -> float:
        return self.omega

    def _check(self, converged: bool) -> None:
        if self.last_residual > self.tol:
            raise RuntimeError(
                f"Jacobi pressure solve stalled: relative residual {self.last_residual:.3e} > "
                f"tol {self.tol:.1e} after {self.last_iterations} sweeps. Use "
                "pressure_solver='direct' or raise poisson_maxiter."
            )


class SORSolver(PoissonSolverBase):
    """Red/black SOR (pressure.py:457-526)."""

    name = "sor"
    warm_startable = True
    _device_kind = "sor"

    def __init__(self, system, tol=DEFAULT_POISSON_TOL, maxiter=DEFAULT_POISSON_MAXITER,
                 omega: float = DEFAULT_SOR_OMEGA, **kw):
        super().__init__(system, tol, maxiter, **kw)
        self.omega = omega
        self._validate_colouring()

    def _validate_colouring(self) -> None:
        # pressure.py:486-496: an active face joining two cells of one colour
        # (odd periodic extent) makes the red/black sweep wrong.
        nx, ny = self.system.shape
        A = self.system.A.tocoo()
        off = A.row != A.col
        r, c = A.row[off], A.col[off]
        ri, rj = np.divmod(r, ny)
        ci, cj = np.divmod(c, ny)
        same = ((ri + rj) % 2) == ((ci + cj) % 2)
        for label, colour in (("red", 0), ("black", 1)):
            if np.any(same & (((ri + rj) % 2) == colour) & (A.data[off] != 0)):
                raise ValueError(
                    f"red/black ordering is invalid for this grid ({label} cells are "
                    "coupled to each other). This happens with an odd number of cells "
                    "on a periodic axis; use an even nx/ny or pressure_solver='direct'."
                )

    def _omega(self) -> float:
        return self.omega

    def _check(self, converged: bool) -> None:
        if self.last_residual > self.tol:
            raise RuntimeError(
                f"SOR pressure solve stalled: relative residual {self.last_residual:.3e} > "
                f"tol {self.tol:.1e} after {self.last_iterations} sweeps. Use "
                "pressure_solver='direct' or raise poisson_maxiter."
            )


class MultigridSolverBase(PoissonSolverBase):
    """Shared hierarchy setup (pressure.py:529-542)."""

    def __init__(self, system, tol=DEFAULT_POISSON_TOL, maxiter=DEFAULT_POISSON_MAXITER,
                 smooth_sweeps: int = DEFAULT_MG_SWEEPS, **kw):
        super().__init__(system, tol, maxiter, **kw)
        from .multigrid import shared_hierarchy
        self.smooth_sweeps = smooth_sweeps
        self.hierarchy = shared_hierarchy(system.A, system.shape, system.fluid, system.singular,
                                          sweeps=smooth_sweeps,
                                          periodic=(system.periodic_x, system.periodic_y),
                                          coarse_max=self.coarse_max)

    def _levels(self):
        return self.hierarchy.device_levels(self.coarse_max, self.wide_classes)

    def _sweeps(self) -> int:
        return self.smooth_sweeps


class MultigridSolver(MultigridSolverBase):
    """Standalone V-cycles with the stall guard (pressure.py:545-574)."""

    name = "multigrid"
    warm_startable = True
    _device_kind = "multigrid"

    def _check(self, converged: bool) -> None:
        if self.last_residual > self.tol:
            raise RuntimeError(
                f"multigrid pressure solve stalled: relative residual {self.last_residual:.3e} "
                f"> tol {self.tol:.1e} after {self.last_iterations} V-cycles. Use "
                "pressure_solver='direct' or raise poisson_maxiter."
            )


class MultigridCGSolver(MultigridSolverBase, _CGFamily):
    """CG preconditioned by one V-cycle (pressure.py:577-632)."""

    name = "mgcg"
    warm_startable = True
    _device_kind = "mgcg"
    _label = "multigrid-preconditioned CG"


class DirectSolver(MultigridSolverBase, _CGFamily):
    """The ``direct`` selection (pressure.py:330-363) on the GPU.

    There is no sparse LU on the device: the system is solved by
    multigrid-preconditioned CG to ``DIRECT_TOL``, i.e. to round-off level,
    which is what the reference's LU delivers.  ``last_iterations`` reports the
    CG iterations taken instead of the reference's constant 1.
    """

    name = "direct"
    warm_startable = False
    _device_kind = "direct"
    _label = "direct (multigrid-CG) pressure solve"

    def _tol_maxiter(self) -> tuple[float, int]:
        return DIRECT_TOL, DIRECT_MAXITER

    def _check(self, converged: bool) -> None:
        if not converged and self.last_residual > 1e-11:
            raise RuntimeError(
                f"direct (multigrid-CG) pressure solve failed to converge "
                f"(relative residual={self.last_residual:.3e} after "
                f"{self.last_iterations} iterations)"
            )


_SOLVERS: dict[PressureSolver, type[PoissonSolverBase]] = {
    PressureSolver.DIRECT: DirectSolver,
    PressureSolver.CG: CGSolver,
    PressureSolver.JACOBI: JacobiSolver,
    PressureSolver.SOR: SORSolver,
    PressureSolver.MULTIGRID: MultigridSolver,
    PressureSolver.MGCG: MultigridCGSolver,
}


def make_pressure_solver(kind: PressureSolver, system: PoissonSystem,
                         tol: float = DEFAULT_POISSON_TOL,
                         maxiter: int = DEFAULT_POISSON_MAXITER,
                         sor_omega: float = DEFAULT_SOR_OMEGA,
                         mg_sweeps: int = DEFAULT_MG_SWEEPS, *, context=None,
                         precision: str = "fp64",
                         mg_coarse_max: int = DEFAULT_MG_COARSE_MAX) -> PoissonSolverBase:
    """Instantiate the requested solver (pressure.py:645-660)."""
    kind = PressureSolver(kind)
    cls = _SOLVERS[kind]
    kw = {"context": context, "precision": precision, "coarse_max": mg_coarse_max}
    if cls is SORSolver:
        return SORSolver(system, tol, maxiter, omega=sor_omega, **kw)
    if issubclass(cls, MultigridSolverBase):
        return cls(system, tol, maxiter, smooth_sweeps=mg_sweeps, **kw)
    return cls(system, tol, maxiter, **kw)
