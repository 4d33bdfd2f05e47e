"""Wall conditions: host-side description of ``pycfd/core/boundary.py``.

The device kernels (``csrc/boundary.hip``) apply the conditions; this module
holds what the host needs to describe them: the wall index table
(boundary.py:50-80), the per-kind properties (boundary.py:99-333) and the
inlet profiles (boundary.py:205-216), computed with the reference's own
arithmetic so the device receives bit-identical lines.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..config import WALLS, BCKind, BCSpec
from .mesh import StructuredMesh


@dataclass(frozen=True)
class WallIndex:
    """Array indices of one wall on the staggered grid (boundary.py:50-66)."""

    axis: int
    n_face: int
    n_in: int
    n_ghost: int
    t_ghost: int
    t_in: int
    p_ghost: int
    p_in: int
    outward: int


def wall_index(wall: str, mesh: StructuredMesh) -> WallIndex:
    """boundary.py:69-80."""
    nx, ny = mesh.shape
    table = {
        "left": WallIndex(0, 1, 2, 0, 0, 1, 0, 1, -1),
        "right": WallIndex(0, nx + 1, nx, nx + 2, nx + 1, nx, nx + 1, nx, +1),
        "bottom": WallIndex(1, 1, 2, 0, 0, 1, 0, 1, -1),
        "top": WallIndex(1, ny + 1, ny, ny + 2, ny + 1, ny, ny + 1, ny, +1),
    }
    try:
        return table[wall]
    except KeyError:
        raise ValueError(f"unknown wall {wall!r}, expected one of {WALLS}") from None


class WallCondition:
    """Properties of one wall's condition that the host and tests consult."""

    def __init__(self, wall: str, spec: BCSpec) -> None:
        if wall not in WALLS:
            raise ValueError(f"unknown wall {wall!r}, expected one of {WALLS}")
        self.wall = wall
        self.spec = spec
        self.kind = spec.kind
        self.pressure_kind = {BCKind.PRESSURE_OUTLET: "dirichlet",
                              BCKind.PERIODIC: "periodic"}.get(spec.kind, "neumann")
        self.derives_normal = spec.kind is BCKind.OUTLET
        self.prescribes_normal = spec.kind is not BCKind.PRESSURE_OUTLET

    @property
    def p_ref(self) -> float:
        return self.spec.p_ref

    def inlet_line(self, mesh: StructuredMesh) -> np.ndarray | None:
        """Full normal-velocity line of an inlet wall (boundary.py:205-216)."""
        if self.kind is not BCKind.INLET:
            return None
        wi = wall_index(self.wall, mesh)
        length = (mesh.ny + 2) if wi.axis == 0 else (mesh.nx + 2)
        speed = -wi.outward * self.spec.velocity
        if self.spec.profile == "uniform":
            return np.full(length, speed, dtype=float)
        coord, extent = (mesh.yc, mesh.ly) if wi.axis == 0 else (mesh.xc, mesh.lx)
        prof = 4.0 * coord * (extent - coord) / extent ** 2
        full = np.zeros(length)
        full[1:1 + coord.size] = speed * prof
        return full

    def __repr__(self) -> str:
        return f"{self.kind.value}(wall={self.wall!r})"


class BoundaryManager:
    """The four wall conditions of a case (boundary.py:359-372)."""

    def __init__(self, boundary_config: dict[str, BCSpec], mesh: StructuredMesh) -> None:
        self.mesh = mesh
        self.conditions = {w: WallCondition(w, boundary_config[w]) for w in WALLS}
        self.periodic_x = self.conditions["left"].kind is BCKind.PERIODIC
        self.periodic_y = self.conditions["bottom"].kind is BCKind.PERIODIC
        self._outlets = [w for w, bc in self.conditions.items() if bc.kind is BCKind.OUTLET]
        self._inlets = [w for w, bc in self.conditions.items() if bc.kind is BCKind.INLET]

    def pressure_boundary_kinds(self) -> dict[str, str]:
        return {w: bc.pressure_kind for w, bc in self.conditions.items()}

    def dirichlet_pressure(self) -> dict[str, float]:
        return {w: bc.p_ref for w, bc in self.conditions.items()
                if bc.pressure_kind == "dirichlet"}

    def __repr__(self) -> str:
        inner = ", ".join(f"{w}={bc.kind.value}" for w, bc in self.conditions.items())
        return f"BoundaryManager({inner})"
