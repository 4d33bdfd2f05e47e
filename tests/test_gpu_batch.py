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


def test_batched_instances_equal_lone_runs(gpu_available):
    from pycfd_amd.batch import SimulationBatch, cavity_batch_configs
    from pycfd_amd.physics.incompressible import Simulation
    cfgs = cavity_batch_configs(count=7, seed=5, sizes=(24, 32, 40, 48),
                                pressure_solver="mgcg", poisson_tol=1e-12)
    batch = SimulationBatch(cfgs, streams=3)
    res = batch.run(max_steps=6)
    assert res.steps == [6] * 7 and res.mlups > 0
    for cfg, sim in zip(cfgs, batch.sims):
        lone = Simulation(cfg)
        lone.run(max_steps=6)
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
