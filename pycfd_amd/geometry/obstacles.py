"""Immersed bodies as cell masks (the parts of ``pycfd/geometry/obstacles.py``
the projection step consumes: the circle mask of the cylinder case and the
multi-body grouping that per-body forces need)."""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

from ..core.mesh import StructuredMesh

DEFAULT_SUBSAMPLES = 8
SOLID_FRACTION_THRESHOLD = 0.5


@dataclass
class Obstacle:
    """A solid body: boolean ``(nx, ny)`` mask plus its volume fractions (obstacles.py:55-84)."""

    mask: np.ndarray
    fraction: np.ndarray
    characteristic_length: float
    name: str = "obstacle"

    @property
    def area(self) -> float:
        return float(self.fraction.sum())


def cell_solid_fraction(mesh: StructuredMesh, inside, subsamples: int = DEFAULT_SUBSAMPLES):
    """Mid-point supersampled solid fraction per cell (obstacles.py:87-103)."""
    if subsamples < 1:
        raise ValueError(f"subsamples must be >= 1, got {subsamples}")
    offs = (np.arange(subsamples) + 0.5) / subsamples
    xs = (mesh.xf[:-1, None] + offs[None, :] * mesh.dx_cells[:, None]).ravel()
    ys = (mesh.yf[:-1, None] + offs[None, :] * mesh.dy_cells[:, None]).ravel()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    hits = inside(X, Y).astype(float)
    return hits.reshape(mesh.nx, subsamples, mesh.ny, subsamples).mean(axis=(1, 3))


def circle_mask(mesh: StructuredMesh, center: tuple[float, float], radius: float,
                subsamples: int = DEFAULT_SUBSAMPLES, name: str = "cylinder") -> Obstacle:
    """Circular body, solid where more than half a cell is covered (obstacles.py:106-135)."""
    if radius <= 0:
        raise ValueError(f"radius must be positive, got {radius}")
    cx, cy = center
    if not (0 <= cx <= mesh.lx and 0 <= cy <= mesh.ly):
        raise ValueError(f"circle centre {center} lies outside the domain")
    frac = cell_solid_fraction(
        mesh, lambda X, Y: (X - cx) ** 2 + (Y - cy) ** 2 <= radius ** 2, subsamples)
    mask = frac > SOLID_FRACTION_THRESHOLD
    if not mask.any():
        raise ValueError(
            f"circle of radius {radius:g} does not cover a single cell on a "
            f"{mesh.nx}x{mesh.ny} grid")
    return Obstacle(mask, frac, 2.0 * radius, name)


def rectangle_mask(mesh: StructuredMesh, lower: tuple[float, float],
                   upper: tuple[float, float], subsamples: int = DEFAULT_SUBSAMPLES,
                   name: str = "rectangle") -> Obstacle:
    (x0, y0), (x1, y1) = lower, upper
    frac = cell_solid_fraction(
        mesh, lambda X, Y: (X >= x0) & (X <= x1) & (Y >= y0) & (Y <= y1), subsamples)
    mask = frac > SOLID_FRACTION_THRESHOLD
    return Obstacle(mask, frac, float(y1 - y0), name)


@dataclass
class ObstacleGroup:
    """Several disjoint bodies; forces are reported per body (obstacles.py:445-614)."""

    members: tuple[Obstacle, ...]

    def __post_init__(self) -> None:
        if not self.members:
            raise ValueError("an obstacle group needs at least one body")
        total = np.zeros_like(self.members[0].mask, dtype=int)
        for b in self.members:
            total += b.mask.astype(int)
        if np.any(total > 1):
            raise ValueError("bodies of a group must not overlap")

    def __len__(self) -> int:
        return len(self.members)

    @property
    def mask(self) -> np.ndarray:
        out = np.zeros_like(self.members[0].mask, dtype=bool)
        for b in self.members:
            out |= b.mask
        return out

    @property
    def masks(self) -> tuple[np.ndarray, ...]:
        return tuple(b.mask for b in self.members)

    @property
    def names(self) -> tuple[str, ...]:
        return tuple(b.name for b in self.members)

    @property
    def name(self) -> str:
        return " + ".join(self.names)

    @property
    def characteristic_length(self) -> float:
        return max(b.characteristic_length for b in self.members)


def as_obstacle_group(obstacle) -> ObstacleGroup | None:
    """Normalise ``None`` / one body / a sequence / a group (obstacles.py:617-638)."""
    if obstacle is None:
        return None
    if isinstance(obstacle, ObstacleGroup):
        return obstacle
    if isinstance(obstacle, Obstacle):
        return ObstacleGroup((obstacle,))
    if isinstance(obstacle, np.ndarray):
        m = np.asarray(obstacle, dtype=bool)
        return ObstacleGroup((Obstacle(m, m.astype(float), 1.0, "mask"),))
    if isinstance(obstacle, Sequence):
        return ObstacleGroup(tuple(obstacle))
    raise TypeError(f"cannot interpret {type(obstacle).__name__} as an obstacle")
