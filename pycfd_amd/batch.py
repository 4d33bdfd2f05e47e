"""Many independent simulations advanced together on one GPU.

The batched-cavity workload: ``n`` lid-driven cavities whose grid sizes are
drawn uniformly from a set and whose Reynolds numbers are log-uniform in a
range, each set up exactly as ``lid_driven_cavity.build`` does
(pycfd/cases/lid_driven_cavity.py:59-85: unit square, lid u = 1, dt ceiling
0.01, CFL 0.4, the reference's default pressure solver) and each advanced
a fixed number of steps.

Each instance is an ordinary ``Simulation`` with its own device context on
its own HIP stream; ``SimulationBatch.run`` drives the streams from host
threads (the library releases the interpreter while the device works), so
the instances' kernels overlap on the device's hardware queues instead of
running one instance at a time.  An instance's arithmetic is exactly that of
the same simulation run alone.
"""

from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .config import BCKind, BCSpec, SimulationConfig

CAVITY_SIZES = (64, 96, 128, 192, 256, 384, 512)


def cavity_config(n: int, re: float, pressure_solver: str = "direct", **overrides) -> SimulationConfig:
    """lid_driven_cavity.build's configuration for an n x n cavity at ``re``."""
    overrides.setdefault("name", f"cavity_{n}_Re{re:.4g}")
    overrides.setdefault("t_end", 1.0e9)
    return SimulationConfig(
        nx=n, ny=n, lx=1.0, ly=1.0, re=float(re), u_ref=1.0, l_ref=1.0, dt=0.01, cfl_max=0.4,
        use_les=False, pressure_solver=pressure_solver,
        boundary_config={"left": BCSpec(BCKind.NO_SLIP), "right": BCSpec(BCKind.NO_SLIP),
                         "bottom": BCSpec(BCKind.NO_SLIP),
                         "top": BCSpec(BCKind.MOVING_WALL, velocity=1.0)},
        **overrides)


def cavity_batch_configs(count: int = 96, seed: int = 2024, sizes=CAVITY_SIZES,
                         re_range=(100.0, 1000.0), **overrides) -> list[SimulationConfig]:
    """``count`` cavities: sizes uniform over ``sizes``, Re log-uniform over
    ``re_range``, from ``numpy.random.default_rng(seed)``."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(sizes), size=count)
    res = np.exp(rng.uniform(np.log(re_range[0]), np.log(re_range[1]), size=count))
    return [cavity_config(int(n), float(re), **overrides) for n, re in zip(ns, res)]


@dataclass
class BatchResult:
    steps: list[int]
    wall_time: float
    cells: list[int]
    setup_time: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def cell_steps(self) -> int:
        return int(sum(c * s for c, s in zip(self.cells, self.steps)))

    @property
    def mlups(self) -> float:
        return self.cell_steps / self.wall_time / 1e6 if self.wall_time > 0 else float("nan")


class SimulationBatch:
    """Independent simulations, one HIP stream each (shared round-robin when
    there are more instances than ``streams``).  ``warm_start_order`` sets
    every instance's ``ProjectionSolver.warm_start_order`` (0: the
    reference's start from the last pressure)."""

    def __init__(self, configs, obstacles=None, inits=None, streams: int = 8,
                 warm_start_order: int = 0) -> None:
        import torch

        from .physics.incompressible import Simulation
        if not configs:
            raise ValueError("a batch needs at least one configuration")
        n = len(configs)
        obstacles = list(obstacles) if obstacles is not None else [None] * n
        inits = list(inits) if inits is not None else [(None, None)] * n
        if len(obstacles) != n or len(inits) != n:
            raise ValueError("obstacles and inits must match the configurations")
        self.n_streams = max(1, min(int(streams), n))
        self.streams = [torch.cuda.Stream() for _ in range(self.n_streams)]
        self.sims = []
        t0 = time.perf_counter()
        for k, cfg in enumerate(configs):
            with torch.cuda.stream(self.streams[k % self.n_streams]):
                u0, v0 = inits[k]
                sim = Simulation(cfg, obstacle=obstacles[k], u_init=u0, v_init=v0)
                if warm_start_order:
                    sim.solver.warm_start_order = warm_start_order
                self.sims.append(sim)
        torch.cuda.synchronize()
        self.setup_time = time.perf_counter() - t0

    def __len__(self) -> int:
        return len(self.sims)

    def run(self, max_steps: int, threads: int | None = None) -> BatchResult:
        """Advance every instance ``max_steps`` steps.  One host thread per
        stream steps that stream's instances round-robin, one step each in
        turn; an error in any instance is raised here."""
        import torch
        threads = self.n_streams if threads is None else max(1, min(threads, self.n_streams))
        groups = [[k for k in range(len(self.sims)) if k % self.n_streams == s]
                  for s in range(self.n_streams)]
        workers = [[] for _ in range(threads)]
        for s, g in enumerate(groups):
            workers[s % threads].append((s, g))
        done = [0] * len(self.sims)
        errors: list[BaseException] = []

        def work(assigned):
            try:
                pending = [k for _, g in assigned for k in g]
                stream_of = {k: self.streams[s] for s, g in assigned for k in g}
                for _ in range(max_steps):
                    for k in pending:
                        with torch.cuda.stream(stream_of[k]):
                            res = self.sims[k].run(max_steps=1)
                        done[k] += res.steps
            except BaseException as e:  # noqa: BLE001 -- re-raised in the caller
                errors.append(e)

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ts = [threading.Thread(target=work, args=(w,)) for w in workers if w]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if errors:
            raise errors[0]
        cells = [s.mesh.nx * s.mesh.ny for s in self.sims]
        return BatchResult(steps=done, wall_time=wall, cells=cells, setup_time=self.setup_time)


__all__ = ["BatchResult", "CAVITY_SIZES", "SimulationBatch", "cavity_batch_configs",
           "cavity_config"]
