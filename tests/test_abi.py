"""The C-ABI library: loads, exports exactly what include/pycfd_hip.h
declares, and its struct layouts agree with the ctypes binding.  No compute
call is made; on a machine without a GPU the library must refuse to create a
context rather than fall back to anything."""

from __future__ import annotations

import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from pycfd_amd import _lib

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "pycfd_hip.h"


def declared() -> set[str]:
    text = HEADER.read_text()
    return set(re.findall(r"^int\s+(pcfd_\w+)\s*\(", text, flags=re.M))


def exported() -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T pcfd_" in line}


def test_library_built_and_loads():
    assert _lib.LIB_PATH.exists(), "run __graft_entry__.build() first"
    _lib.load()


def test_exports_every_declared_symbol():
    assert declared() == exported()
    assert set(_lib.EXPORTED) == declared()


def test_clock_slots_match_header():
    """The live clock's slot table: the binding names as many kinds as the
    header counts (the C side files its kinds in the same order)."""
    from pycfd_amd.core.engine import DeviceContext as Engine
    text = HEADER.read_text()
    kinds = int(re.search(r"#define PCFD_CLOCK_KINDS (\d+)", text).group(1))
    levels = int(re.search(r"#define PCFD_CLOCK_LEVELS (\d+)", text).group(1))
    assert len(Engine.CLOCK_KINDS) == kinds
    assert Engine.CLOCK_LEVELS == levels
    src = (ROOT / "pycfd_amd" / "csrc" / "poisson.hpp").read_text()
    enum = re.search(r"enum ClockKind : int \{(.*?)\};", src, flags=re.S).group(1)
    names = [x.split("=")[0].strip() for x in enum.split(",") if x.strip()]
    assert names[-1] == "kClockKinds" and len(names) - 1 == kinds


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(f"""
#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
#define S(T) printf(#T " %zu\\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\\n", offsetof(T, f));
int main(void) {{
  S(pcfd_case_desc) O(pcfd_case_desc, time_scheme) O(pcfd_case_desc, inlet_line)
  O(pcfd_case_desc, solid) O(pcfd_case_desc, n_bodies)
  S(pcfd_solver_desc) O(pcfd_solver_desc, precision)
  S(pcfd_mg_level_desc) O(pcfd_mg_level_desc, cls) O(pcfd_mg_level_desc, weight)
  O(pcfd_mg_level_desc, coarse_inv)
  S(pcfd_solve_info) O(pcfd_solve_info, residual) O(pcfd_solve_info, status)
  S(pcfd_mg_operator) O(pcfd_mg_operator, x0) O(pcfd_mg_operator, every_row)
  S(pcfd_poisson_assembly) O(pcfd_poisson_assembly, cx_face) O(pcfd_poisson_assembly, p_ref)
  S(pcfd_mg_built_level) O(pcfd_mg_built_level, nnz) O(pcfd_mg_built_level, weight)
  return 0;
}}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in
               subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines())
    py = {
        "pcfd_case_desc": _lib.CaseDesc, "pcfd_solver_desc": _lib.SolverDesc,
        "pcfd_mg_level_desc": _lib.MGLevelDesc, "pcfd_solve_info": _lib.SolveInfo,
        "pcfd_mg_operator": _lib.MGOperator, "pcfd_mg_built_level": _lib.MGBuiltLevel,
        "pcfd_poisson_assembly": _lib.PoissonAssembly,
    }
    for key, val in got.items():
        if "." in key:
            t, f = key.split(".")
            assert getattr(py[t], f).offset == int(val), key
        else:
            assert C.sizeof(py[key]) == int(val), key


def _desc(nx, ny, kinds, solid=None):
    d = _lib.CaseDesc()
    d.nx, d.ny, d.dx, d.dy = nx, ny, 1.0 / nx, 1.0 / ny
    for w, k in enumerate(kinds):
        d.bc_kind[w] = _lib.BC_CODES[k]
    keep = None
    if solid is not None:
        keep = np.ascontiguousarray(solid, dtype=np.uint8)
        d.solid = keep.ctypes.data_as(C.POINTER(C.c_uint8))
    return d, keep


@pytest.mark.parametrize("kinds", [("no_slip",) * 4, ("periodic",) * 4,
                                   ("inlet", "pressure_outlet", "symmetry", "symmetry")])
def test_device_operator_flags_match_assembled_matrix(kinds):
    """The flag byte the device applies encodes exactly the host CSR operator."""
    from pycfd_amd.core.mesh import StructuredMesh
    from pycfd_amd.core.pressure import assemble_poisson_matrix
    nx, ny = 13, 10
    solid = np.zeros((nx, ny), bool)
    solid[4:7, 3:6] = True
    solid[10, 9] = True
    d, keep = _desc(nx, ny, kinds, solid)
    flags = np.zeros(nx * ny, np.uint8)
    n = C.c_int32(0)
    assert _lib.load().pcfd_poisson_flags(C.byref(d), flags.ctypes.data_as(C.POINTER(C.c_uint8)),
                                          C.byref(n)) == 0
    mesh = StructuredMesh(nx, ny)
    dirichlet = {"right": 0.0} if "pressure_outlet" in kinds else None
    S = assemble_poisson_matrix(mesh, kinds[0] == "periodic", kinds[2] == "periodic",
                                solid, dirichlet)
    A = S.A.tocoo()
    assert np.array_equal((flags & 1).astype(bool), S.fluid)
    assert n.value == int(S.fluid.sum())
    bits = {-ny: 2, ny: 4, -1: 8, 1: 16}
    rebuilt = np.zeros(nx * ny, np.uint8)
    for r, c in zip(A.row, A.col):
        if r == c:
            continue
        ri, rj = divmod(r, ny)
        ci, cj = divmod(c, ny)
        di = (ci - ri + nx // 2) % nx - nx // 2 if kinds[0] == "periodic" else ci - ri
        dj = (cj - rj + ny // 2) % ny - ny // 2 if kinds[2] == "periodic" else cj - rj
        rebuilt[r] |= bits[di * ny + dj]
    assert np.array_equal(flags & 0x1E, rebuilt)


def test_refuses_without_device():
    """With no HIP device the engine reports it instead of computing anything."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except Exception:
        pass
    d, _ = _desc(8, 8, ("no_slip",) * 4)
    h = C.c_void_p(0)
    rc = _lib.load().pcfd_ctx_create(C.byref(d), None, C.byref(h))
    assert rc == _lib.PCFD_ERR_NO_DEVICE and not h.value
    assert "device" in _lib.last_error()


def test_product_has_no_oracle_or_cpu_path():
    """The package never imports the oracle."""
    pkg = ROOT / "pycfd_amd"
    for f in pkg.rglob("*.py"):
        text = f.read_text()
        assert "oracle" not in text.replace("the oracle", ""), f
