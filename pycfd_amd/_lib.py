"""ctypes binding of ``libpycfd_hip.so`` (declared in ``include/pycfd_hip.h``).

The library is the only compute path: there is no CPU fallback.  Loading
fails loudly when the shared object is missing, and every status code the
library returns is turned into the exception type the reference raises for
the same condition (pycfd/core/pressure.py:393-397, timestepper.py:47-48).
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("PYCFD_HIP_LIB", _HERE / "libpycfd_hip.so"))

PCFD_OK = 0
PCFD_ERR_HIP = 1
PCFD_ERR_NOT_CONVERGED = 2
PCFD_ERR_STALLED = 3
PCFD_ERR_NONFINITE = 4
PCFD_ERR_BAD_ARG = 5
PCFD_ERR_UNSUPPORTED = 6
PCFD_ERR_NO_DEVICE = 7

BC_CODES = {
    "no_slip": 0, "moving_wall": 1, "inlet": 2, "outlet": 3,
    "pressure_outlet": 4, "periodic": 5, "symmetry": 6,
}
SOLVER_CODES = {"direct": 0, "cg": 1, "jacobi": 2, "sor": 3, "multigrid": 4, "mgcg": 5}
SCHEME_CODES = {"euler": 0, "rk2": 1, "rk3": 2}
MASK_ZERO, MASK_RESET, MASK_ACCUMULATE = 0, 1, 2

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i8p = C.POINTER(C.c_int8)


class CaseDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double),
        ("nu", C.c_double),
        ("body_fx", C.c_double), ("body_fy", C.c_double),
        ("time_scheme", C.c_int32),
        ("bc_kind", C.c_int32 * 4),
        ("bc_velocity", C.c_double * 4),
        ("bc_p_ref", C.c_double * 4),
        ("inlet_line", _dp * 4),
        ("solid", _u8p),
        ("body_id", _i8p),
        ("n_bodies", C.c_int32),
        ("use_les", C.c_int32),
        ("les_coeff", C.c_double),
        ("dx_cells", _dp), ("dy_cells", _dp),
        ("area_mean", C.c_double),
        ("les_cs", C.c_double),
        ("soft_ib", C.c_int32),
    ]


class SolverDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("tol", C.c_double), ("maxiter", C.c_int32),
        ("omega", C.c_double), ("sweeps", C.c_int32), ("precision", C.c_int32),
        ("fused_levels", C.c_int32), ("tile_tx", C.c_int32), ("tile_ty", C.c_int32),
        ("coarse_chain", C.c_int32), ("stream_rows", C.c_int32),
        ("row_passes", C.c_int32),
    ]


class MGLevelDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("fx", C.c_int32), ("fy", C.c_int32),
        ("n_offsets", C.c_int32), ("di", _i8p), ("dj", _i8p),
        ("n_classes", C.c_int32), ("cls", C.POINTER(C.c_uint16)), ("table", _dp),
        ("class_dinv", _dp), ("class_active", _u8p), ("active", _u8p),
        ("weight", C.c_double), ("coarse_n", C.c_int32), ("coarse_inv", _dp),
        ("cls32", C.POINTER(C.c_uint32)),
    ]


class PoissonAssembly(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("periodic_x", C.c_int32), ("periodic_y", C.c_int32),
        ("solid", _u8p), ("cx", C.c_double), ("cy", C.c_double),
        ("cx_face", _dp), ("cy_face", _dp),
        ("dirichlet", C.c_int32 * 4), ("p_ref", C.c_double * 4),
    ]


class MGOperator(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("indptr", C.POINTER(C.c_int64)), ("indices", C.POINTER(C.c_int32)),
        ("data", _dp), ("active", _u8p), ("x0", _dp),
        ("singular", C.c_int32), ("periodic_x", C.c_int32), ("periodic_y", C.c_int32),
        ("threads", C.c_int32), ("coarse_max", C.c_int32), ("every_row", C.c_int32),
    ]


class MGBuiltLevel(C.Structure):
    _fields_ = [
        ("n_levels", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
        ("fx", C.c_int32), ("fy", C.c_int32), ("pin", C.c_int32), ("n_offsets", C.c_int32),
        ("nnz", C.c_int64), ("n_classes", C.c_int64), ("weight", C.c_double),
    ]


class SolveInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("residual", C.c_double),
        ("converged", C.c_int32), ("status", C.c_int32),
    ]


_SIGNATURES = {
    "pcfd_ctx_create": [C.POINTER(CaseDesc), C.c_void_p, C.POINTER(C.c_void_p)],
    "pcfd_ctx_destroy": [C.c_void_p],
    "pcfd_last_error": [C.c_char_p, C.c_size_t],
    "pcfd_device_count": [C.POINTER(C.c_int32)],
    "pcfd_synchronize": [C.c_void_p],
    "pcfd_poisson_flags": [C.POINTER(CaseDesc), _u8p, C.POINTER(C.c_int32)],
    "pcfd_poisson_assemble": [C.POINTER(PoissonAssembly)] + [C.c_void_p] * 5,
    "pcfd_mg_build": [C.POINTER(MGOperator), C.POINTER(C.c_void_p)],
    "pcfd_mg_free": [C.c_void_p],
    "pcfd_mg_level_info": [C.c_void_p, C.c_int32, C.POINTER(MGBuiltLevel)],
    "pcfd_mg_level_copy": [C.c_void_p, C.c_int32] + [C.c_void_p] * 11,
    "pcfd_set_pressure_solver": [C.c_void_p, C.POINTER(SolverDesc), C.POINTER(MGLevelDesc),
                                 C.c_int32],
    "pcfd_momentum_rhs": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                          C.c_void_p],
    "pcfd_divergence": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_apply_velocity_bc": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32],
    "pcfd_mass_balance": [C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_mask_obstacle": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int32],
    "pcfd_wrap_faces": [C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_apply_pressure_bc": [C.c_void_p, C.c_void_p],
    "pcfd_pressure_solve": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                            C.POINTER(SolveInfo)],
    "pcfd_project": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                     C.POINTER(SolveInfo)],
    "pcfd_initialize": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(SolveInfo)],
    "pcfd_step": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.POINTER(SolveInfo)],
    "pcfd_reset_guess": [C.c_void_p],
    "pcfd_set_step_graphs": [C.c_void_p, C.c_int32],
    "pcfd_set_warm_start": [C.c_void_p, C.c_int32],
    "pcfd_soft_face_masks": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_soft_tape_size": [C.c_void_p, C.POINTER(C.c_int64)],
    "pcfd_soft_step": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_soft_step_adjoint": [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                               C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_soft_cell_gradient": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_step_graph_stats": [C.c_void_p, C.POINTER(C.c_int32)],
    "pcfd_diagnostics": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                         C.c_void_p, _dp],
    "pcfd_get_force": [C.c_void_p, _dp],
    "pcfd_diagnostics_noalloc": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p],
    "pcfd_poisson_apply": [C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_mg_apply": [C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_time_kernel": [C.c_void_p, C.c_int32, C.c_int32, _dp, _dp],
    "pcfd_kernel_clock_start": [C.c_void_p, C.c_int32],
    "pcfd_kernel_clock_read": [C.c_void_p, C.POINTER(C.c_int32), _dp, _dp],
    "pcfd_mg_plan": [C.c_void_p, C.POINTER(C.c_int32), C.c_int32],
    "pcfd_pressure_guess": [C.c_void_p, C.c_void_p, C.c_int32],
    "pcfd_eddy_viscosity": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "pcfd_momentum_rhs_variable": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                   C.c_void_p, C.c_void_p],
}

EXPORTED = tuple(_SIGNATURES)


class HipEngineError(RuntimeError):
    """A HIP runtime failure inside the engine."""


_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the library once; raise ``ImportError`` if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m pycfd_amd.build` "
            "(the HIP engine has no CPU fallback)"
        )
    # torch ships its own HIP runtime; loading it first makes the library's
    # libamdhip64 dependency resolve to that one, so the process holds a
    # single runtime (a second one, loaded first, leaves torch's unable to
    # see the device)
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - torch is a dependency
        pass
    lib = C.CDLL(str(LIB_PATH))
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    load().pcfd_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    """Map a status code to the reference's exception types."""
    if rc == PCFD_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc in (PCFD_ERR_BAD_ARG,):
        raise ValueError(msg)
    if rc == PCFD_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == PCFD_ERR_NO_DEVICE:
        raise RuntimeError(f"no HIP device: {msg}")
    raise HipEngineError(msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def device_count() -> int:
    n = C.c_int32(0)
    load().pcfd_device_count(C.byref(n))
    return int(n.value)
