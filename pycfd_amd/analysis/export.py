"""Checkpoints: the raw ghosted state plus the configuration, for an exact restart.

The archive layout is the reference's (pycfd/analysis/export.py:214-276), so
either code base reads the other's files: ``np.savez_compressed`` with keys
``u``, ``v``, ``p`` (ghosted FP64 arrays), ``t``, ``step``, ``config`` (the
configuration as JSON) and ``provenance`` (a JSON record).  One key is added:
``p_guess``, the Poisson warm-start pressure ``(nx, ny)``, written when a
solver is passed.  The reference does not persist it, so its iterative
solvers resume from a cold start; with it a resumed run repeats the
uninterrupted one bit for bit.  Loaders that do not know the key ignore it.
"""

from __future__ import annotations

import json
import platform
from datetime import datetime, timezone
from pathlib import Path

import numpy as np

from ..config import SimulationConfig
from ..core.fields import FlowField
from ..core.mesh import StructuredMesh


def provenance_record(config: SimulationConfig | None = None, extra: dict | None = None) -> dict:
    """What produced a file: generator, time, interpreter, platform, config."""
    from .. import __version__
    record = {
        "generator": "pycfd_amd",
        "pycfd_amd_version": __version__,
        "generated_utc": datetime.now(timezone.utc).isoformat(timespec="seconds"),
        "python": platform.python_version(),
        "platform": platform.platform(),
    }
    if config is not None:
        record["config"] = config.to_dict()
    if extra:
        record.update(extra)
    return record


def units_record() -> dict:
    """The ``units`` block of every written record: the solver's own
    non-dimensional units (export.py:69-77, no rescaling on this path)."""
    return {"system": "solver", "note": "non-dimensional; u_ref = l_ref = 1"}


def with_units(record: dict) -> dict:
    """A copy of ``record`` carrying the units block, also for a record the
    caller supplied (export.py:80-84)."""
    out = dict(record)
    out["units"] = units_record()
    return out


def save_checkpoint(fields: FlowField, config: SimulationConfig, path: str | Path,
                    provenance: dict | None = None, *, solver=None) -> Path:
    """Write ``fields`` and ``config`` to ``path`` (``.npz`` appended if missing).

    With ``solver`` (a ``ProjectionSolver``), its warm-start pressure is kept
    too, so an iterative pressure solve resumes exactly where it stopped.
    """
    path = Path(path)
    if path.suffix != ".npz":
        path = path.with_suffix(".npz")
    path.parent.mkdir(parents=True, exist_ok=True)
    record = with_units(provenance_record(config) if provenance is None else provenance)
    arrays = dict(u=fields.u, v=fields.v, p=fields.p, t=np.array(fields.t),
                  step=np.array(fields.step), config=np.array(config.to_json()),
                  provenance=np.array(json.dumps(record)))
    guess = solver.pressure_guess() if solver is not None else None
    if guess is not None:
        arrays["p_guess"] = guess
    np.savez_compressed(path, **arrays)
    path.with_suffix(".provenance.json").write_text(json.dumps(record, indent=2))
    return path


def checkpoint_provenance(path: str | Path) -> dict | None:
    """The provenance embedded in a checkpoint, or ``None`` if it has none."""
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"checkpoint not found: {path}")
    with np.load(path, allow_pickle=False) as npz:
        if "provenance" not in npz.files:
            return None
        return json.loads(str(npz["provenance"]))


def load_checkpoint(path: str | Path) -> tuple[FlowField, SimulationConfig]:
    """Restore a state written by :func:`save_checkpoint` (export.py:253-276).

    The fields come back as host arrays and move to the GPU on first use.
    """
    fields, config, _ = load_checkpoint_with_guess(path)
    return fields, config


def load_checkpoint_with_guess(path: str | Path):
    """``(fields, config, p_guess or None)``."""
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"checkpoint not found: {path}")
    with np.load(path, allow_pickle=False) as npz:
        config = SimulationConfig.from_dict(json.loads(str(npz["config"])))
        mesh = StructuredMesh.from_config(config)
        u, v, p = npz["u"], npz["v"], npz["p"]
        for arr, expected, label in ((u, mesh.u_shape, "u"), (v, mesh.v_shape, "v"),
                                     (p, mesh.p_shape, "p")):
            if arr.shape != tuple(expected):
                raise ValueError(
                    f"checkpoint {label} has shape {arr.shape}, expected {tuple(expected)}; "
                    "the file does not match its stored configuration")
        fields = FlowField(mesh, u, v, p, float(npz["t"]), int(npz["step"]))
        guess = np.array(npz["p_guess"]) if "p_guess" in npz.files else None
    if guess is not None and guess.shape != tuple(mesh.shape):
        raise ValueError(f"checkpoint p_guess has shape {guess.shape}, expected {mesh.shape}")
    return fields, config, guess


__all__ = ["checkpoint_provenance", "load_checkpoint", "load_checkpoint_with_guess",
           "provenance_record", "save_checkpoint", "units_record", "with_units"]
