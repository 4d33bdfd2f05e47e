"""Metric reductions the cases report (host side, on fetched fields).

Follows ``pycfd/analysis/postprocess.py`` (vorticity :21-40, kinetic energy
:64-72, force coefficients :104-113, Strouhal :116-142, centre-line profiles
and probes :148-221) and ``pycfd/analysis/timeseries.py:216-291`` for the
time average that defines a mean force coefficient.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def vorticity_corner(fields) -> np.ndarray:
    """``dv/dx - du/dy`` at the corners, each derivative over the centre-to-centre span."""
    nx, ny = fields.mesh.shape
    m = fields.mesh.metrics
    u, v = fields.u, fields.v
    dvdx = (v[1:nx + 2, 1:ny + 2] - v[0:nx + 1, 1:ny + 2]) / m.hxu
    dudy = (u[1:nx + 2, 1:ny + 2] - u[1:nx + 2, 0:ny + 1]) / m.hyv
    return dvdx - dudy


def vorticity(fields) -> np.ndarray:
    w = vorticity_corner(fields)
    return 0.25 * (w[:-1, :-1] + w[1:, :-1] + w[:-1, 1:] + w[1:, 1:])


def kinetic_energy(fields, mask: np.ndarray | None = None) -> float:
    uc, vc = fields.cell_velocities()
    e = 0.5 * (uc ** 2 + vc ** 2)
    if mask is not None:
        e = np.where(mask, 0.0, e)
    return fields.mesh.integrate(e)


def force_coefficients(force, u_ref: float, l_ref: float,
                       density: float = 1.0) -> tuple[float, float]:
    """``C = F / (0.5 rho U^2 L)`` (postprocess.py:104-113)."""
    if u_ref <= 0 or l_ref <= 0:
        raise ValueError("u_ref and l_ref must be positive to form a coefficient")
    q = 0.5 * density * u_ref ** 2 * l_ref
    return force[0] / q, force[1] / q


def strouhal_number(times, signal, l_ref: float, u_ref: float) -> float:
    """Dominant frequency of the mean-removed, Hann-windowed, uniformly
    resampled signal, as ``f L / U`` (postprocess.py:116-142)."""
    times = np.asarray(times, dtype=float)
    signal = np.asarray(signal, dtype=float)
    if times.size < 16:
        return float("nan")
    uniform_t = np.linspace(times[0], times[-1], times.size)
    s = np.interp(uniform_t, times, signal)
    s = s - s.mean()
    dt = uniform_t[1] - uniform_t[0]
    if dt <= 0:
        return float("nan")
    spectrum = np.abs(np.fft.rfft(s * np.hanning(s.size)))
    freqs = np.fft.rfftfreq(s.size, dt)
    spectrum[0] = 0.0
    peak = int(np.argmax(spectrum))
    if peak == 0:
        return float("nan")
    return float(freqs[peak] * l_ref / u_ref)


def late_time_mean(times, values, fraction: float = 0.5) -> float:
    """Mean over ``t > fraction * t_end`` (timeseries.py:255-259)."""
    t = np.asarray(times, dtype=float)
    x = np.asarray(values, dtype=float)
    keep = t > fraction * t[-1]
    return float(x[keep].mean()) if keep.any() else float(x.mean())


def _interp_axis(arr, coords, target, axis):
    i = int(np.clip(np.searchsorted(coords, target) - 1, 0, coords.size - 2))
    w = (target - coords[i]) / (coords[i + 1] - coords[i])
    lo = arr[i] if axis == 0 else arr[:, i]
    hi = arr[i + 1] if axis == 0 else arr[:, i + 1]
    return (1.0 - w) * lo + w * hi


def centerline_profiles(fields):
    """u on x = lx/2 and v on y = ly/2 (postprocess.py:148-161)."""
    mesh = fields.mesh
    uc, vc = fields.cell_velocities()
    u_line = _interp_axis(uc, mesh.xc, 0.5 * mesh.lx, axis=0)
    v_line = _interp_axis(vc, mesh.yc, 0.5 * mesh.ly, axis=1)
    return mesh.yc, u_line, mesh.xc, v_line


def _bilinear(arr, xs, ys, x, y) -> float:
    i = int(np.clip(np.searchsorted(xs, x) - 1, 0, xs.size - 2))
    j = int(np.clip(np.searchsorted(ys, y) - 1, 0, ys.size - 2))
    tx = np.clip((x - xs[i]) / (xs[i + 1] - xs[i]), 0.0, 1.0)
    ty = np.clip((y - ys[j]) / (ys[j + 1] - ys[j]), 0.0, 1.0)
    return ((1 - tx) * (1 - ty) * arr[i, j] + tx * (1 - ty) * arr[i + 1, j]
            + (1 - tx) * ty * arr[i, j + 1] + tx * ty * arr[i + 1, j + 1])


def sample_at(fields, x: float, y: float) -> dict[str, float]:
    mesh = fields.mesh
    uc, vc = fields.cell_velocities()
    return {name: float(_bilinear(arr, mesh.xc, mesh.yc, x, y))
            for name, arr in (("u", uc), ("v", vc), ("p", fields.p_phys))}


@dataclass
class Probe:
    """Point time series (postprocess.py:195-221)."""

    x: float
    y: float
    name: str = "probe"
    t: list = field(default_factory=list)
    u: list = field(default_factory=list)
    v: list = field(default_factory=list)
    p: list = field(default_factory=list)

    def record(self, fields) -> None:
        s = sample_at(fields, self.x, self.y)
        self.t.append(fields.t)
        self.u.append(s["u"])
        self.v.append(s["v"])
        self.p.append(s["p"])

    def as_arrays(self) -> dict[str, np.ndarray]:
        return {k: np.asarray(getattr(self, k), dtype=float) for k in ("t", "u", "v", "p")}
