"""Flow state ``(u, v, p)`` resident in HBM, with the reference's host view.

``FlowField`` mirrors ``pycfd/core/fields.py:18-96``.  The arrays live on the
GPU as FP64 torch tensors in the reference's ghosted C-order layout; ``u``,
``v``, ``p`` hand back read-only NumPy copies (fetched once and cached), so
code written against the reference's ``FlowField`` reads the same numbers.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import StructuredMesh


def _torch():
    import torch
    return torch


def device_zeros(shape) -> "object":
    torch = _torch()
    return torch.zeros(shape, dtype=torch.float64, device="cuda")


def device_empty(shape) -> "object":
    """Uninitialised device array, for outputs a kernel writes in full."""
    torch = _torch()
    return torch.empty(shape, dtype=torch.float64, device="cuda")


def to_device(a) -> "object":
    torch = _torch()
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def to_host(t) -> np.ndarray:
    torch = _torch()
    if isinstance(t, torch.Tensor):
        return t.detach().to("cpu").numpy()
    return np.asarray(t)


@dataclass
class FlowField:
    """The staggered state plus simulation clock (fields.py:18-39).

    ``u``, ``v``, ``p`` may be given as device tensors or as host arrays (a
    restored checkpoint); host arrays are uploaded on first device use.
    """

    mesh: StructuredMesh
    u_src: object
    v_src: object
    p_src: object
    t: float = 0.0
    step: int = 0
    diagnostics: dict = field(default_factory=dict)
    _host: dict = field(default_factory=dict, repr=False)

    def __post_init__(self) -> None:
        for name in ("u", "v", "p"):
            src = getattr(self, name + "_src")
            if isinstance(src, np.ndarray):
                arr = np.array(src, dtype=np.float64, order="C")
                arr.flags.writeable = False
                self._host[name] = arr
                setattr(self, name + "_src", None)

    @classmethod
    def zeros(cls, mesh: StructuredMesh) -> "FlowField":
        return cls(mesh, device_zeros(mesh.u_shape), device_zeros(mesh.v_shape),
                   device_zeros(mesh.p_shape))

    def _dev(self, name: str):
        src = getattr(self, name + "_src")
        if src is None:
            src = to_device(self._host[name])
            setattr(self, name + "_src", src)
        return src

    @property
    def u_dev(self):
        return self._dev("u")

    @property
    def v_dev(self):
        return self._dev("v")

    @property
    def p_dev(self):
        return self._dev("p")

    def copy(self) -> "FlowField":
        return FlowField(self.mesh, self.u_dev.clone(), self.v_dev.clone(), self.p_dev.clone(),
                         self.t, self.step, dict(self.diagnostics))

    def invalidate(self) -> None:
        """Drop cached host copies after the device arrays were changed in place."""
        for name in ("u", "v", "p"):
            self._dev(name)
        self._host.clear()

    def _view(self, name: str) -> np.ndarray:
        arr = self._host.get(name)
        if arr is None:
            arr = to_host(getattr(self, name + "_src")).copy()
            arr.flags.writeable = False
            self._host[name] = arr
        return arr

    @property
    def u(self) -> np.ndarray:
        return self._view("u")

    @property
    def v(self) -> np.ndarray:
        return self._view("v")

    @property
    def p(self) -> np.ndarray:
        return self._view("p")

    @property
    def u_phys(self) -> np.ndarray:
        nx, ny = self.mesh.shape
        return self.u[1:nx + 2, 1:ny + 1]

    @property
    def v_phys(self) -> np.ndarray:
        nx, ny = self.mesh.shape
        return self.v[1:nx + 1, 1:ny + 2]

    @property
    def p_phys(self) -> np.ndarray:
        nx, ny = self.mesh.shape
        return self.p[1:nx + 1, 1:ny + 1]

    def cell_velocities(self) -> tuple[np.ndarray, np.ndarray]:
        """Velocity at cell centres (fields.py:70-79)."""
        nx, ny = self.mesh.shape
        u, v = self.u, self.v
        uc = 0.5 * (u[1:nx + 1, 1:ny + 1] + u[2:nx + 2, 1:ny + 1])
        vc = 0.5 * (v[1:nx + 1, 1:ny + 1] + v[1:nx + 1, 2:ny + 2])
        return uc, vc

    def speed(self) -> np.ndarray:
        uc, vc = self.cell_velocities()
        return np.hypot(uc, vc)

    def max_velocity(self) -> tuple[float, float]:
        """``(max|u|, max|v|)`` over the physical faces (fields.py:86-88)."""
        d = self.diagnostics
        if "umax" in d and "vmax" in d:
            return d["umax"], d["vmax"]
        return float(np.abs(self.u_phys).max()), float(np.abs(self.v_phys).max())

    def is_finite(self) -> bool:
        d = self.diagnostics
        if "nonfinite" in d:
            return d["nonfinite"] == 0
        return bool(np.all(np.isfinite(self.u)) and np.all(np.isfinite(self.v))
                    and np.all(np.isfinite(self.p)))
