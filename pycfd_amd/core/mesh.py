"""Structured MAC mesh: coordinates and ghosted array shapes (host side).

Index convention and coordinate arithmetic follow ``pycfd/core/mesh.py``
(layout :14-31, widths :269-311, faces and centres :353-375, shapes
:483-508) so that initial conditions and body masks evaluated on this mesh are
bit-identical to the reference's.  A uniform mesh hands the device its scalar
``dx``, ``dy``; a stretched one its cell widths, from which the library forms
the :class:`MeshMetrics` arrays with the same elementwise arithmetic.
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


class MeshMetrics:
    """The spacings the staggered operators divide by (mesh.py:122-233).

    Plain floats on a uniform mesh, broadcastable arrays (``(n, 1)`` for x,
    ``(1, n)`` for y) on a stretched one: ``hx, hy`` cell widths, ``hxu, hyv``
    centre-to-centre spans across each face (ghost cells mirror their
    neighbour, so ``hxu[0] == hx[0]``), ``hx_ext, hy_ext`` widths with those
    ghosts, ``wx_corner, wy_corner`` weights of the upper node at a corner,
    ``cell_areas``, ``min_hx, min_hy``.
    """

    __slots__ = ("hx", "hy", "hxu", "hyv", "hx_ext", "hy_ext", "wx_corner", "wy_corner",
                 "cell_areas", "min_hx", "min_hy", "uniform_area")

    def __init__(self, mesh: "StructuredMesh") -> None:
        dxc, dyc = mesh.dx_cells, mesh.dy_cells
        hx_ext = np.concatenate(([dxc[0]], dxc, [dxc[-1]]))
        hy_ext = np.concatenate(([dyc[0]], dyc, [dyc[-1]]))
        hxu = 0.5 * (hx_ext[:-1] + hx_ext[1:])
        hyv = 0.5 * (hy_ext[:-1] + hy_ext[1:])
        self.cell_areas = dxc[:, None] * dyc[None, :]
        self.min_hx = float(dxc.min())
        self.min_hy = float(dyc.min())
        self.uniform_area = float(dxc[0] * dyc[0]) if mesh.is_uniform else None
        if mesh.is_uniform:
            self.hx = self.hxu = self.hx_ext = float(dxc[0])
            self.hy = self.hyv = self.hy_ext = float(dyc[0])
            self.wx_corner = self.wy_corner = 0.5
        else:
            self.hx, self.hy = dxc[:, None], dyc[None, :]
            self.hxu, self.hyv = hxu[:, None], hyv[None, :]
            self.hx_ext, self.hy_ext = hx_ext[:, None], hy_ext[None, :]
            self.wx_corner = (0.5 * hx_ext[:-1] / hxu)[:, None]
            self.wy_corner = (0.5 * hy_ext[:-1] / hyv)[None, :]

    def integrate(self, field: np.ndarray) -> float:
        """Domain integral of a cell-centred ``(nx, ny)`` field (mesh.py:213-223)."""
        if self.uniform_area is not None:
            return float(np.sum(field)) * self.uniform_area
        return float(np.sum(field * self.cell_areas))

    @staticmethod
    def _slice(value, lo: int, hi: int, axis: int):
        """``value[lo:hi]`` along ``axis``, or the scalar unchanged (mesh.py:225-233)."""
        if np.isscalar(value):
            return value
        return value[lo:hi, :] if axis == 0 else value[:, lo:hi]


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
        self._metrics: MeshMetrics | None = None

    @classmethod
    def from_config(cls, cfg) -> "StructuredMesh":
        return cls(cfg.nx, cfg.ny, cfg.lx, cfg.ly, cfg.stretch_x, cfg.stretch_y,
                   cfg.cluster_x, cfg.cluster_y)

    @property
    def is_uniform(self) -> bool:
        return self._is_uniform

    @property
    def metrics(self) -> MeshMetrics:
        if self._metrics is None:
            self._metrics = MeshMetrics(self)
        return self._metrics

    @property
    def area_mean(self) -> float:
        """Mean cell area, the Poisson scaling of a stretched mesh (pressure.py:197)."""
        return float(self.metrics.cell_areas.mean())

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
        """Domain integral of a cell-centred field (mesh.py:213-223)."""
        return self.metrics.integrate(field)

    def __repr__(self) -> str:
        if not self._is_uniform:
            return (f"StructuredMesh({self.nx}x{self.ny} cells, domain "
                    f"[0,{self.lx:g}]x[0,{self.ly:g}], stretched: dx in "
                    f"[{self.dx_cells.min():.6g}, {self.dx_cells.max():.6g}], dy in "
                    f"[{self.dy_cells.min():.6g}, {self.dy_cells.max():.6g}])")
        return (f"StructuredMesh({self.nx}x{self.ny} cells, domain "
                f"[0,{self.lx:g}]x[0,{self.ly:g}], dx={self._dx:.6g}, dy={self._dy:.6g})")
