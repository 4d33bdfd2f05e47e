"""Build libpycfd_hip.so for gfx950 from the HIP sources in ``csrc/``.

The library is built in-tree (next to this file) so that it travels with the
source tree.  ``hipcc`` cross-compiles the device code without a GPU present.

Flags:
  -ffp-contract=off   no fused multiply-add: the stencils reproduce the
                      reference's NumPy arithmetic bit for bit.  The fused
                      multigrid kernels (vcycle.hip), a preconditioner whose
                      rounding no result depends on bit for bit, contract
                      multiply-adds (-ffp-contract=fast).
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keep the compiler from forming packed FP32 arithmetic.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libpycfd_hip.so"
BUILD = HERE / "_build"
ARCH = os.environ.get("PYCFD_HIP_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value",
]
CONTRACT = {"vcycle": "-ffp-contract=fast"}


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, CONTRACT.get(src.stem, "-ffp-contract=off"), "-c", str(src),
           "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{proc.stdout}\n{proc.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every ``csrc/*.hip`` and link ``libpycfd_hip.so``."""
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    workers = max(1, min(len(srcs), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(workers) as pool:
        objs = list(pool.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
