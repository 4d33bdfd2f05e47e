"""Structured MAC mesh: coordinates and ghosted array shapes (host side).

Index convention and coordinate arithmetic follow ``pycfd/core/mesh.py``
(layout :14-31, widths :269-311, faces and centres :353-375, shapes
:483-508) so that initial conditions and body masks evaluated on this mesh are
bit-identical to the reference's.  Device kernels only need ``dx``, ``dy``.
"""

from __future__ import annotations

import numpy as np

from ..config import _normalise_cluster


class NonUniformMeshError(RuntimeError):
    """A uniform-spacing operator was handed a stretched mesh (mesh.py:114)."""


_UNIFORMITY_RTOL = 1.0e-12


def _geometric_widths(length: float, n: int, ratio: float, cluster: str = "low") -> np.ndarray:
    """``n`` widths growing by ``ratio`` and summing to ``length`` (mesh.py:269-311)."""
    if ratio <= 0:
        raise ValueError(f"stretch ratio must be positive, got {ratio}")
    if abs(ratio - 1.0) < 1.0e-14:
        return np.full(n, length / n, dtype=float)
    if cluster == "low":
        w0 = length * (ratio - 1.0) / (ratio ** n - 1.0)
        return w0 * ratio ** np.arange(n, dtype=float)
    half, odd = divmod(n, 2)
    if cluster == "walls":
        run = ratio ** np.arange(half, dtype=float)
        w0 = length / (2.0 * run.sum() + odd * ratio ** half)
        run = w0 * run
        middle = [w0 * ratio ** half] if odd else []
        return np.concatenate([run, middle, run[::-1]])
    if odd:
        out = ratio ** np.arange(1, half + 1, dtype=float)
        w0 = length / (1.0 + 2.0 * out.sum())
        out = w0 * out
        return np.concatenate([out[::-1], [w0], out])
    run = ratio ** np.arange(half, dtype=float)
    run = (length / (2.0 * run.sum())) * run
    return np.concatenate([run[::-1], run])


class StructuredMesh:
    """Cartesian mesh on ``[0, lx] x [0, ly]`` with ``nx x ny`` cells."""

    def __init__(self, nx: int, ny: int, lx: float = 1.0, ly: float = 1.0,
                 stretch_x: float = 1.0, stretch_y: float = 1.0,
                 cluster_x: str = "low", cluster_y: str = "low") -> None:
        if nx < 4 or ny < 4:
            raise ValueError(f"mesh must be at least 4x4 cells, got {nx}x{ny}")
        if lx <= 0 or ly <= 0:
            raise ValueError(f"domain extents must be positive, got {lx}x{ly}")
        self.nx, self.ny = int(nx), int(ny)
        self.lx, self.ly = float(lx), float(ly)
        self.stretch_x, self.stretch_y = float(stretch_x), float(stretch_y)
        self.cluster_x = _normalise_cluster(cluster_x)
        self.cluster_y = _normalise_cluster(cluster_y)
        self.dx_cells = _geometric_widths(self.lx, self.nx, self.stretch_x, self.cluster_x)
        self.dy_cells = _geometric_widths(self.ly, self.ny, self.stretch_y, self.cluster_y)
        self.xf = np.concatenate(([0.0], np.cumsum(self.dx_cells)))
        self.yf = np.concatenate(([0.0], np.cumsum(self.dy_cells)))
        self.xf[-1] = self.lx
        self.yf[-1] = self.ly
        self.xc = 0.5 * (self.xf[:-1] + self.xf[1:])
        self.yc = 0.5 * (self.yf[:-1] + self.yf[1:])
        self.stretched_x = not bool(
            np.allclose(self.dx_cells, self.dx_cells[0], rtol=_UNIFORMITY_RTOL))
        self.stretched_y = not bool(
            np.allclose(self.dy_cells, self.dy_cells[0], rtol=_UNIFORMITY_RTOL))
        self._is_uniform = not (self.stretched_x or self.stretched_y)
        self._dx = float(self.dx_cells[0])
        self._dy = float(self.dy_cells[0])

    @classmethod
    def from_config(cls, cfg) -> "StructuredMesh":
        return cls(cfg.nx, cfg.ny, cfg.lx, cfg.ly, cfg.stretch_x, cfg.stretch_y,
                   cfg.cluster_x, cfg.cluster_y)

    @property
    def is_uniform(self) -> bool:
        return self._is_uniform

    def require_uniform(self, who: str = "operator") -> None:
        if not self._is_uniform:
            raise NonUniformMeshError(
                f"{who} is discretised for uniform spacing but the mesh is "
                f"stretched (stretch_x={self.stretch_x}, stretch_y={self.stretch_y}). "
                "Use stretch_x = stretch_y = 1.0."
            )

    @property
    def dx(self) -> float:
        self.require_uniform("StructuredMesh.dx")
        return self._dx

    @property
    def dy(self) -> float:
        self.require_uniform("StructuredMesh.dy")
        return self._dy

    @property
    def cell_area(self) -> float:
        self.require_uniform("StructuredMesh.cell_area")
        return self._dx * self._dy

    @property
    def shape(self) -> tuple[int, int]:
        return (self.nx, self.ny)

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny

    def cell_centers(self):
        return self.xc, self.yc

    def face_centers(self):
        return self.xf, self.yf

    def u_coords(self):
        return self.xf, self.yc

    def v_coords(self):
        return self.xc, self.yf

    def cell_center_grid(self):
        return np.meshgrid(self.xc, self.yc, indexing="ij")

    def u_grid(self):
        return np.meshgrid(self.xf, self.yc, indexing="ij")

    def v_grid(self):
        return np.meshgrid(self.xc, self.yf, indexing="ij")

    @property
    def u_shape(self) -> tuple[int, int]:
        return (self.nx + 3, self.ny + 2)

    @property
    def v_shape(self) -> tuple[int, int]:
        return (self.nx + 2, self.ny + 3)

    @property
    def p_shape(self) -> tuple[int, int]:
        return (self.nx + 2, self.ny + 2)

    def zeros_u(self) -> np.ndarray:
        return np.zeros(self.u_shape)

    def zeros_v(self) -> np.ndarray:
        return np.zeros(self.v_shape)

    def zeros_p(self) -> np.ndarray:
        return np.zeros(self.p_shape)

    def integrate(self, field: np.ndarray) -> float:
        """Domain integral of a cell-centred field (mesh.py:213-223, uniform)."""
        return float(np.sum(field)) * (self._dx * self._dy)

    def __repr__(self) -> str:
        return (f"StructuredMesh({self.nx}x{self.ny} cells, domain "
                f"[0,{self.lx:g}]x[0,{self.ly:g}], dx={self._dx:.6g}, dy={self._dy:.6g})")
