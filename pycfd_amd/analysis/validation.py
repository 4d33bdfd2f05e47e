"""Closed-form solutions and published reference data used by the cases.

* Taylor-Green vortex on a 2*pi-periodic box (pycfd/analysis/validation.py:47-65).
* Ghia, Ghia & Shin, J. Comput. Phys. 48 (1982) 387-411, Tables I and II:
  lid-driven-cavity centre-line velocities (the published table, used by
  pycfd/analysis/validation.py:80-142; one Re=400 entry the reference marks
  unverified is stored as NaN and skipped, as there).
* RMS / max error norms (validation.py:160-190).
"""

from __future__ import annotations

import numpy as np


def taylor_green(x, y, t: float, nu: float, component: str) -> np.ndarray:
    decay = np.exp(-2.0 * nu * t)
    if component == "u":
        return -np.cos(x) * np.sin(y) * decay
    if component == "v":
        return np.sin(x) * np.cos(y) * decay
    if component == "p":
        return -0.25 * (np.cos(2.0 * x) + np.cos(2.0 * y)) * decay ** 2
    raise ValueError(f"component must be 'u', 'v' or 'p', got {component!r}")


# Sample positions along the vertical (y) and horizontal (x) centre lines.
_GHIA_Y = (0.0000, 0.0547, 0.0625, 0.0703, 0.1016, 0.1719, 0.2813, 0.4531, 0.5000,
           0.6172, 0.7344, 0.8516, 0.9531, 0.9609, 0.9688, 0.9766, 1.0000)
_GHIA_X = (0.0000, 0.0625, 0.0703, 0.0781, 0.0938, 0.1563, 0.2266, 0.2344, 0.5000,
           0.8047, 0.8594, 0.9063, 0.9453, 0.9531, 0.9609, 0.9688, 1.0000)
_GHIA = {
    100: {
        "u": (0.00000, -0.03717, -0.04192, -0.04775, -0.06434, -0.10150, -0.15662,
              -0.21090, -0.20581, -0.13641, 0.00332, 0.23151, 0.68717, 0.73722,
              0.78871, 0.84123, 1.00000),
        "v": (0.00000, 0.09233, 0.10091, 0.10890, 0.12317, 0.16077, 0.17507, 0.17527,
              0.05454, -0.24533, -0.22445, -0.16914, -0.10313, -0.08864, -0.07391,
              -0.05906, 0.00000),
    },
    400: {
        "u": (0.00000, -0.08186, -0.09266, -0.10338, -0.14612, -0.24299, -0.32726,
              -0.17119, -0.11477, 0.02135, 0.16256, 0.29093, 0.55892, 0.61756,
              0.68439, 0.75837, 1.00000),
        "v": (0.00000, 0.18360, 0.19713, 0.20920, 0.22965, 0.28124, 0.30203, 0.30174,
              0.05186, -0.38598, -0.44993, float("nan"), -0.22847, -0.19254, -0.15663,
              -0.12146, 0.00000),
    },
    1000: {
        "u": (0.00000, -0.18109, -0.20196, -0.22220, -0.29730, -0.38289, -0.27805,
              -0.10648, -0.06080, 0.05702, 0.18719, 0.33304, 0.46604, 0.51117,
              0.57492, 0.65928, 1.00000),
        "v": (0.00000, 0.27485, 0.29012, 0.30353, 0.32627, 0.37095, 0.33075, 0.32235,
              0.02526, -0.31966, -0.42665, -0.51550, -0.39188, -0.33714, -0.27669,
              -0.21388, 0.00000),
    },
}
GHIA_REYNOLDS = tuple(sorted(_GHIA))


def poiseuille_profile(y, height: float, u_max: float) -> np.ndarray:
    """Plane Poiseuille profile ``u_max 4 y (h - y) / h^2`` (validation.py:26-31)."""
    y = np.asarray(y, dtype=float)
    return u_max * 4.0 * y * (height - y) / height ** 2


def convergence_orders(resolutions, errors) -> list[float]:
    """``log(e1 / e2) / log(N2 / N1)`` per refinement pair (validation.py:223-244)."""
    out = []
    for k in range(len(resolutions) - 1):
        e1, e2 = float(errors[k]), float(errors[k + 1])
        n1, n2 = int(resolutions[k]), int(resolutions[k + 1])
        out.append(float(np.log(e1 / e2) / np.log(n2 / n1)) if e1 > 0 and e2 > 0
                   else float("nan"))
    return out


def ghia_reference(re: int) -> dict[str, np.ndarray]:
    re = int(re)
    if re not in _GHIA:
        raise ValueError(f"no Ghia et al. reference data for Re={re}; "
                         f"available: {list(GHIA_REYNOLDS)}")
    return {"y": np.array(_GHIA_Y), "u": np.array(_GHIA[re]["u"]),
            "x": np.array(_GHIA_X), "v": np.array(_GHIA[re]["v"])}


def _aligned(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    if a.shape != b.shape:
        raise ValueError(f"shape mismatch: {a.shape} vs {b.shape}")
    keep = np.isfinite(a) & np.isfinite(b)
    return a[keep], b[keep]


def l2_error(numerical, reference, relative: bool = False) -> float:
    a, b = _aligned(numerical, reference)
    err = float(np.sqrt(np.mean((a - b) ** 2)))
    if relative:
        scale = float(np.sqrt(np.mean(b ** 2)))
        return err / scale if scale > 0 else err
    return err


def linf_error(numerical, reference, relative: bool = False) -> float:
    a, b = _aligned(numerical, reference)
    err = float(np.abs(a - b).max())
    if relative:
        scale = float(np.abs(b).max())
        return err / scale if scale > 0 else err
    return err


def ghia_errors(fields, re: float = 100.0) -> dict[str, float]:
    """Centre-line L2 and L-infinity errors against Ghia et al. at ``re``
    (lid_driven_cavity.py:88-117)."""
    from .postprocess import centerline_profiles
    y, u_line, x, v_line = centerline_profiles(fields)
    ref = ghia_reference(int(round(re)))
    u_at = np.interp(ref["y"], y, u_line)
    v_at = np.interp(ref["x"], x, v_line)
    return {"ghia_u_L2": l2_error(u_at, ref["u"]), "ghia_u_Linf": linf_error(u_at, ref["u"]),
            "ghia_v_L2": l2_error(v_at, ref["v"]), "ghia_v_Linf": linf_error(v_at, ref["v"])}
