"""A program that leaves its device contexts open at exit.

It runs a cavity with the cooperative chain mode of the deepest multigrid
levels and the live kernel clock on, then returns without closing anything:
the library's exit hook must close every context before the HIP runtime
shuts down, so the process exits with status 0 (also under a profiler).
"""

from __future__ import annotations

import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def main() -> int:
    import torch

    from pycfd_amd.physics.incompressible import Simulation
    from tests.cases_def import product_config
    torch.cuda.set_device(0)
    cfg, body = product_config("cavity", pressure_solver="mgcg", poisson_tol=1e-12,
                               mg_coarse_max=40)
    sim = Simulation(cfg, obstacle=body)
    ps = sim.solver.pressure_solver
    ps.coarse_chain = 1
    ps.configure()
    # the chain is kept when the device accepts its cooperative launch (it
    # may refuse while another process, such as a test runner, holds it)
    modes = [p["mode"] for p in sim.solver.ctx.mg_plan()]
    sim.solver.ctx.kernel_clock_start(max_launches=100000)
    sim.run(max_steps=3)
    clocked = sim.solver.ctx.kernel_clock_read()
    assert clocked and all(e["launches"] > 0 for e in clocked)
    # a second context, also left open
    other = Simulation(cfg, obstacle=body)
    other.run(max_steps=1)
    torch.cuda.synchronize()
    print(f"teardown child: modes {modes}, {len(clocked)} clocked kernels; exiting with "
          "contexts open", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
