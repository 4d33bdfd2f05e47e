"""Smagorinsky sub-grid-scale closure on the GPU (pycfd/physics/turbulence.py).

``nu_t = (Cs Delta)^2 |S|`` with ``Delta = sqrt(dx dy)`` and
``|S| = sqrt(2 S11^2 + 2 S22^2 + 4 S12^2)``: normal rates at cell centres, the
shear rate at corners averaged to the centres, copied into the ghost ring.  The
momentum right-hand side then switches to the stress-divergence form of the
viscous term (solver.py:467-499).  Both run in the HIP stencil kernels; this
class keeps the reference's interface for callers and tests.
"""

from __future__ import annotations

import numpy as np

from ..core.mesh import StructuredMesh


class SmagorinskyModel:
    """turbulence.py:25-113, evaluated on the GPU.

    On a stretched mesh the filter width is per cell, ``sqrt(hx hy)``, and
    ``coeff`` is the ghosted ``(nx+2, ny+2)`` array the reference keeps
    (turbulence.py:47-58); the device forms the same values itself.
    """

    def __init__(self, mesh: StructuredMesh, cs: float = 0.17) -> None:
        if cs < 0:
            raise ValueError(f"Smagorinsky constant must be non-negative, got {cs}")
        self.mesh = mesh
        self.cs = float(cs)
        if mesh.is_uniform:
            self.delta: float | np.ndarray = float(np.sqrt(mesh.dx * mesh.dy))
            self._coeff: float | np.ndarray = (self.cs * self.delta) ** 2
        else:
            delta = np.sqrt(mesh.dx_cells[:, None] * mesh.dy_cells[None, :])
            self.delta = delta
            nx, ny = mesh.shape
            coeff = np.empty((nx + 2, ny + 2))
            coeff[1:nx + 1, 1:ny + 1] = (self.cs * delta) ** 2
            coeff[0, :], coeff[nx + 1, :] = coeff[1, :], coeff[nx, :]
            coeff[:, 0], coeff[:, ny + 1] = coeff[:, 1], coeff[:, ny]
            self._coeff = coeff
        self.last_nu_t_max = 0.0
        self._ctx = None
        self._nu = None

    @property
    def coeff(self):
        """``(Cs Delta)^2``: a float, or the ghosted per-cell array on a stretched mesh."""
        return self._coeff

    @property
    def device_coeff(self) -> float:
        """The scalar the device takes (unused on a stretched mesh, which sends ``cs``)."""
        return float(self._coeff) if np.isscalar(self._coeff) else 0.0

    def bind(self, ctx, nu: float) -> None:
        """Attach the device context whose kernels evaluate the closure."""
        self._ctx = ctx
        self._nu = float(nu)

    def eddy_viscosity(self, u, v, nu_molecular: float):
        """``(nu_c (nx+2, ny+2), nu_corner (nx+1, ny+1))`` (turbulence.py:101-113)."""
        import torch

        from ..core.fields import to_device, to_host
        if self._ctx is None:
            raise RuntimeError("SmagorinskyModel is not bound to a device context")
        if float(nu_molecular) != self._nu:
            raise ValueError(f"the device closure was set up for nu = {self._nu}, "
                             f"got {nu_molecular}")
        nx, ny = self.mesh.shape
        nu_c = torch.empty((nx + 2, ny + 2), dtype=torch.float64, device="cuda")
        nu_k = torch.empty((nx + 1, ny + 1), dtype=torch.float64, device="cuda")
        self._ctx.eddy_viscosity(to_device(u), to_device(v), nu_c, nu_k)
        nu_c_h, nu_k_h = to_host(nu_c), to_host(nu_k)
        self.last_nu_t_max = float((nu_c_h - self._nu).max())
        return nu_c_h, nu_k_h

    def __repr__(self) -> str:
        if np.isscalar(self.delta):
            return f"SmagorinskyModel(cs={self.cs:g}, delta={self.delta:.4g})"
        return (f"SmagorinskyModel(cs={self.cs:g}, delta in "
                f"[{self.delta.min():.4g}, {self.delta.max():.4g}])")


__all__ = ["SmagorinskyModel"]
