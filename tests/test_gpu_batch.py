"""Batched cavities: independent instances advanced together on one GPU.

Each instance runs on its own stream, driven by its own host thread; the
overlap must not change any instance's arithmetic (bit-identical to the
same simulation run alone), and each instance is the reference's cavity
(compared with the oracle).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("warm", [0, 4])
def test_batched_instances_equal_lone_runs(gpu_available, warm):
    """``warm``: every instance's solves started from the extrapolation in
    time (its history lives in the instance's own context)."""
    from pycfd_amd.batch import SimulationBatch, cavity_batch_configs
    from pycfd_amd.physics.incompressible import Simulation
    cfgs = cavity_batch_configs(count=7, seed=5, sizes=(24, 32, 40, 48),
                                pressure_solver="mgcg", poisson_tol=1e-12)
    batch = SimulationBatch(cfgs, streams=3, warm_start_order=warm)
    steps = 8 if warm else 6
    res = batch.run(max_steps=steps)
    assert res.steps == [steps] * 7 and res.mlups > 0
    for cfg, sim in zip(cfgs, batch.sims):
        assert sim.solver.warm_start_order == warm
        lone = Simulation(cfg)
        lone.solver.warm_start_order = warm
        lone.run(max_steps=steps)
        assert np.array_equal(lone.fields.u, sim.fields.u), cfg.name
        assert np.array_equal(lone.fields.v, sim.fields.v), cfg.name
        assert np.array_equal(lone.fields.p, sim.fields.p), cfg.name


def test_batched_instances_match_oracle(gpu_available):
    """Two instances of different size and Re, the reference's direct solver
    (the cavity case's default): velocity within 1e-10 relative, pressure
    within 1e-8 of the oracle after 8 steps."""
    from pycfd_amd.batch import SimulationBatch, cavity_config
    cfgs = [cavity_config(32, 157.0), cavity_config(40, 820.0)]
    batch = SimulationBatch(cfgs, streams=2)
    batch.run(max_steps=8)
    for cfg, sim in zip(cfgs, batch.sims):
        case = O.Case(nx=cfg.nx, ny=cfg.ny, re=cfg.re, dt=0.01, cfl_max=0.4,
                      kinds=("no_slip", "no_slip", "no_slip", "moving_wall"),
                      velocity=(0.0, 0.0, 0.0, 1.0), pressure_solver="direct", tol=1e-12)
        u, v, p, hist, t = O.run(case, 8)
        assert sim.fields.t == pytest.approx(t, rel=1e-12)
        assert np.abs(sim.fields.u - u).max() <= 1e-10
        assert np.abs(sim.fields.v - v).max() <= 1e-10
        assert np.abs(sim.fields.p - p).max() <= 1e-8 * max(np.abs(p).max(), 1.0)
