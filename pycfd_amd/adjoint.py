"""Differentiable drag of a soft-masked cylinder, with a hand-written adjoint.

The reference holds a body with a binary mask (pycfd/core/solver.py:553-599)
and offers no derivatives.  Here the body is a per-cell field
``m = sigmoid(theta)``; the step forces each solvable face by its soft weight
``m_f = 1 - (1 - m_a)(1 - m_b)`` (the binary "touches a solid" made smooth)
and books the removed momentum as the reaction force, so the time-averaged
drag coefficient is a smooth function of ``theta`` (see pcfd_soft_step in
include/pycfd_hip.h).  With ``m`` in {0, 1} the step is the binary one.

``SoftMaskDrag.loss`` runs the forward on the GPU, keeping the state and the
Poisson warm start every ``checkpoint_every`` steps; ``gradient`` walks the
segments backwards: each segment is recomputed from its checkpoint with its
steps' states recorded, and the hand-written adjoint HIP kernels
(csrc/adjoint.hip) carry the adjoint through every step -- transposed wall
conditions, the Jacobian-transpose of the momentum right-hand side, the
transposed projection with the same symmetric multigrid-CG solve, the soft
forcing with its force term -- accumulating dL/d(face weights), which
``pcfd_soft_cell_gradient`` maps to dL/dtheta.  ``soft_mask_drag`` puts it
behind ``torch.autograd.Function``.

The case: a channel with uniform inflow (u = 1), a pressure outlet (p = 0),
slip walls, a cylinder of diameter ``diameter`` at ``center``,
``theta = -sharpness * SDF`` (positive inside), fixed dt (so the time grid
does not depend on theta), starting from the uniform stream.  The Poisson
operator keeps the topology of the cells where the initial ``theta > 0``
(the binary body); faces touching those cells are forced to zero as in the
binary step, so the soft weights, and the gradient, live on the cells of the
fluid band around the body.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .config import BCKind, BCSpec, SimulationConfig
from .core.mesh import StructuredMesh
from .core.solver import ProjectionSolver


def _p(t) -> C.c_void_p:
    return C.c_void_p(0 if t is None else int(t.data_ptr()))


@dataclass
class _Checkpoint:
    u: object
    v: object
    guess: object


class SoftMaskDrag:
    """Time-averaged drag coefficient of a soft-masked cylinder and its gradient."""

    def __init__(self, nx: int = 512, ny: int = 256, lx: float = 16.0, ly: float = 8.0,
                 re: float = 100.0, dt: float = 0.005, steps: int = 200,
                 checkpoint_every: int = 20, center=(4.0, 4.0), diameter: float = 1.0,
                 sharpness: float = 20.0, pressure_solver: str = "mgcg",
                 poisson_tol: float = 1e-13) -> None:
        import torch
        self.steps = int(steps)
        self.every = max(1, int(checkpoint_every))
        self.dt = float(dt)
        self.diameter = float(diameter)
        self.cfg = SimulationConfig(
            nx=nx, ny=ny, lx=lx, ly=ly, re=re, u_ref=1.0, l_ref=diameter, dt=dt,
            adaptive_dt=False, t_end=1e9, pressure_solver=pressure_solver,
            poisson_tol=poisson_tol, poisson_maxiter=20000,
            boundary_config={"left": BCSpec(BCKind.INLET, velocity=1.0),
                             "right": BCSpec(BCKind.PRESSURE_OUTLET, p_ref=0.0),
                             "bottom": BCSpec(BCKind.SYMMETRY), "top": BCSpec(BCKind.SYMMETRY)})
        self.mesh = StructuredMesh.from_config(self.cfg)
        xc = (np.arange(nx) + 0.5) * (lx / nx)
        yc = (np.arange(ny) + 0.5) * (ly / ny)
        X, Y = np.meshgrid(xc, yc, indexing="ij")
        sdf = np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2) - 0.5 * diameter
        self.sdf = sdf
        self.theta0 = -sharpness * sdf
        self.solid = self.theta0 > 0.0
        self.solver = ProjectionSolver(self.cfg, self.mesh, obstacle=self.solid, soft_body=True)
        self.ctx = self.solver.ctx
        self.lib = _lib.load()
        dev = "cuda"
        self._u_shape, self._v_shape = self.mesh.u_shape, self.mesh.v_shape
        self.mf_u = torch.zeros(self._u_shape, dtype=torch.float64, device=dev)
        self.mf_v = torch.zeros(self._v_shape, dtype=torch.float64, device=dev)
        n = C.c_int64(0)
        self._check(self.lib.pcfd_soft_tape_size(self.ctx.handle, C.byref(n)), "pcfd_soft_tape_size")
        self.tape_doubles = int(n.value)
        self.forces = torch.zeros((self.steps, 2), dtype=torch.float64, device=dev)
        # the initial state: the uniform stream with the wall conditions applied
        u0 = torch.zeros(self._u_shape, dtype=torch.float64, device=dev)
        v0 = torch.zeros(self._v_shape, dtype=torch.float64, device=dev)
        u0[1:nx + 2, 1:ny + 1] = 1.0
        self.ctx.apply_velocity_bc(u0, v0, True)
        self.u_init, self.v_init = u0, v0
        self.loss_scale = 1.0 / (0.5 * 1.0 * 1.0 * diameter)   # C = F / (0.5 rho U^2 L)
        self._theta_key = None
        self._checkpoints: list[_Checkpoint] = []
        self.p = torch.zeros(self.mesh.p_shape, dtype=torch.float64, device=dev)

    def _check(self, rc: int, what: str) -> None:
        _lib.check(rc, what)

    def _masks(self, theta):
        import torch
        theta = torch.as_tensor(theta, dtype=torch.float64, device="cuda")
        m = torch.sigmoid(theta).contiguous()
        self._check(self.lib.pcfd_soft_face_masks(self.ctx.handle, _p(m), _p(self.mf_u),
                                                  _p(self.mf_v)), "pcfd_soft_face_masks")
        return m

    def _step(self, u, v, n: int, tape=None):
        import torch
        un = torch.empty(self._u_shape, dtype=torch.float64, device="cuda")
        vn = torch.empty(self._v_shape, dtype=torch.float64, device="cuda")
        self._check(self.lib.pcfd_soft_step(self.ctx.handle, _p(u), _p(v), C.c_double(self.dt),
                                            _p(self.mf_u), _p(self.mf_v), _p(un), _p(vn),
                                            _p(self.p), _p(tape), _p(self.forces[n])),
                    "pcfd_soft_step")
        return un, vn

    def loss(self, theta) -> float:
        """Mean drag coefficient over the steps (the forward; keeps checkpoints)."""
        import torch
        self._masks(theta)
        self.solver.set_pressure_guess(None)
        u, v = self.u_init, self.v_init
        self._checkpoints = []
        for n in range(self.steps):
            if n % self.every == 0:
                g = torch.empty(self.mesh.shape, dtype=torch.float64, device="cuda")
                self.ctx.pressure_guess(g, store=False)
                self._checkpoints.append(_Checkpoint(u.clone(), v.clone(), g))
            u, v = self._step(u, v, n)
        torch.cuda.synchronize()
        self.u_final, self.v_final = u, v
        self._theta_key = self._key(theta)
        return float(self.forces[:, 0].sum().item()) * self.loss_scale / self.steps

    @staticmethod
    def _key(theta):
        a = np.asarray(theta.detach().cpu().numpy() if hasattr(theta, "detach") else theta,
                       dtype=np.float64)
        return a.tobytes()

    def gradient(self, theta):
        """dL/dtheta (nx, ny) by the adjoint, recomputing each segment from its
        checkpoint (the forward runs first unless it just ran for theta)."""
        import torch
        if self._theta_key != self._key(theta):
            self.loss(theta)
        m = self._masks(theta)
        dev = "cuda"
        lam_u = torch.zeros(self._u_shape, dtype=torch.float64, device=dev)
        lam_v = torch.zeros(self._v_shape, dtype=torch.float64, device=dev)
        g_u = torch.zeros(self._u_shape, dtype=torch.float64, device=dev)
        g_v = torch.zeros(self._v_shape, dtype=torch.float64, device=dev)
        dLdF = self.loss_scale / self.steps
        tapes = torch.empty((self.every, self.tape_doubles), dtype=torch.float64, device=dev)
        for s in reversed(range(len(self._checkpoints))):
            cp = self._checkpoints[s]
            n0 = s * self.every
            n1 = min(self.steps, n0 + self.every)
            self.ctx.pressure_guess(cp.guess, store=True)
            u, v = cp.u, cp.v
            for n in range(n0, n1):
                u, v = self._step(u, v, n, tapes[n - n0])
            for n in reversed(range(n0, n1)):
                self._check(self.lib.pcfd_soft_step_adjoint(
                    self.ctx.handle, _p(tapes[n - n0]), C.c_double(self.dt), _p(self.mf_u),
                    _p(self.mf_v), C.c_double(dLdF), _p(lam_u), _p(lam_v), _p(g_u), _p(g_v)),
                    "pcfd_soft_step_adjoint")
        grad = torch.zeros(self.mesh.shape, dtype=torch.float64, device=dev)
        self._check(self.lib.pcfd_soft_cell_gradient(self.ctx.handle, _p(m), _p(g_u), _p(g_v),
                                                     _p(grad)), "pcfd_soft_cell_gradient")
        torch.cuda.synchronize()
        self.lam_u_initial, self.lam_v_initial = lam_u, lam_v
        return grad


def fd_check(model: SoftMaskDrag, theta=None, cells: int = 64, h: float = 1e-6, seed: int = 0,
             band: float = 0.35, progress=None) -> dict:
    """The adjoint gradient against central differences of the loss at
    ``cells`` cells drawn (``numpy.random.default_rng(seed)``) from the fluid
    band ``|SDF| < band`` around the body; returns the per-cell pairs and the
    relative and absolute errors."""
    import time
    theta = model.theta0.copy() if theta is None else np.asarray(theta, dtype=np.float64)
    t0 = time.perf_counter()
    value = model.loss(theta)
    t1 = time.perf_counter()
    g = model.gradient(theta).cpu().numpy()
    t2 = time.perf_counter()
    if progress:
        progress(f"loss {value:.15e}: forward {t1 - t0:.3f} s, adjoint {t2 - t1:.3f} s")
    rng = np.random.default_rng(seed)
    pool = np.argwhere((np.abs(model.sdf) < band) & ~model.solid)
    pick = pool[rng.choice(len(pool), size=min(cells, len(pool)), replace=False)]
    adj, fd = [], []
    for k, (i, j) in enumerate(pick):
        tp = theta.copy()
        tp[i, j] += h
        tm = theta.copy()
        tm[i, j] -= h
        d = (model.loss(tp) - model.loss(tm)) / (2.0 * h)
        adj.append(float(g[i, j]))
        fd.append(d)
        if progress:
            progress(f"cell ({i}, {j}) sdf {model.sdf[i, j]:+.4f}: adjoint {g[i, j]:+.12e} "
                     f"central difference {d:+.12e}")
    adj_a, fd_a = np.asarray(adj), np.asarray(fd)
    rel = np.abs(adj_a - fd_a) / np.maximum(np.abs(fd_a), 1e-300)
    model.loss(theta)  # leave the model's checkpoints at theta
    return {"loss": value, "cells": pick.tolist(), "adjoint": adj, "central_difference": fd,
            "h": h, "max_rel_err": float(rel.max()), "median_rel_err": float(np.median(rel)),
            "max_abs_err": float(np.abs(adj_a - fd_a).max()),
            "forward_s": t1 - t0, "adjoint_s": t2 - t1}


def _autograd_function():
    import torch

    class _SoftMaskDragFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, model):
            value = model.loss(theta)
            ctx.model = model
            ctx.save_for_backward(theta)
            return torch.tensor(value, dtype=torch.float64, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            (theta,) = ctx.saved_tensors
            g = ctx.model.gradient(theta)
            return grad_out * g.to(theta.device), None

    return _SoftMaskDragFn


_FN = None


def soft_mask_drag(theta, model: SoftMaskDrag):
    """The mean drag coefficient as a differentiable function of theta (torch)."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(theta, model)


__all__ = ["SoftMaskDrag", "fd_check", "soft_mask_drag"]


def main(argv=None) -> int:
    """``python -m pycfd_amd.adjoint``: the finite-difference check of the
    adjoint gradient on the soft-masked cylinder, printed per cell."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--dt", type=float, default=0.005)
    ap.add_argument("--checkpoint-every", type=int, default=20)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--h", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--json", default=None, help="write the record here")
    a = ap.parse_args(argv)
    model = SoftMaskDrag(nx=a.nx, ny=a.ny, steps=a.steps, dt=a.dt,
                         checkpoint_every=a.checkpoint_every)
    rec = fd_check(model, cells=a.cells, h=a.h, seed=a.seed,
                   progress=lambda s: print(s, flush=True))
    rec.update(nx=a.nx, ny=a.ny, steps=a.steps, dt=a.dt, checkpoint_every=a.checkpoint_every)
    print(f"max rel err {rec['max_rel_err']:.3e}, median {rec['median_rel_err']:.3e}, "
          f"max abs err {rec['max_abs_err']:.3e} over {len(rec['adjoint'])} cells", flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rec, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
