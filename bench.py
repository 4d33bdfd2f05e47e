"""Benchmark: MLUPS of the projection time step on one MI355X.

Workload: flow past a cylinder, 2048 x 1024 channel (32 x 16 D), Re = 100,
uniform inflow, zero-gradient OUTLET, slip walls, D = 1 body at 4 D, v seeded
with a 1e-3 Gaussian blob; SSP-RK3, adaptive dt (ceiling 0.02, CFL 0.4);
pressure by multigrid-preconditioned CG, its smoothing sweeps (V(s,s), s in
1..3) picked by the in-project tuner on whole steps of the warmed-up run,
each substep's solve started from the quartic extrapolation in time of that
substep's last five solutions (--warm-start 4; 0 starts from the last
pressure, the reference's way; the tolerance and the divergence target are
the same either way, only the iteration count drops).  Both precision paths
run it: plain FP64, and mixed (FP32 V-cycle inside FP64 CG), at the same fixed divergence
target: the relative residual tolerance (2e-11 by default) is chosen so
that max|div u| <= 1e-10 after every step on both paths from the first one
(the start-up steps are the largest: 5.6e-11; 2.5e-11 let steps 2-4 reach
1.08e-10), and the line says whether it held.

One "step" is one full SSP-RK3 time step (three projections) through
Simulation.run, per-step diagnostics included, with fields resident in HBM.
MLUPS = nx * ny * steps / wall seconds / 1e6.

Kernel rooflines come from a second window (up to 10 steps) right after
the timed one, with the engine's live kernel clock on: every kernel launch of
the steps (predictors, boundary passes, right-hand side, solve start and
final check, scatter, correction, stage combination, diagnostics) and of the
real pressure solves records start/stop events on the solver's stream
(hipExtLaunchKernelGGL), filed by kernel and multigrid level.  `roofline` is
the bandwidth-bound one with the largest share of that window.  The clock is
off in the MLUPS window.  HBM traffic per launch is read from the committed rocprofv3 counter
summary when one exists for the kernel (profiles/pmc_traffic_*.md).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
DIV_TARGET = 1.0e-10
TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_cylinder2048x1024.md"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


# the step's kernels outside the pressure solve's iterations (one clock slot
# each); k_walls and k_boundary_pass touch O(nx + ny) values: no fraction
STEP_KERNELS = ("k_cg_update", "k_predictor", "k_predictor_comb", "k_walls", "k_boundary_pass",
                "k_prepare", "k_init", "k_check", "k_scatter", "k_correct_ghosts", "k_combine",
                "k_diagnostics", "k_warm_guess")


def kernel_name(entry: dict, precision: str, plan: list[dict]) -> str:
    """Name of a clocked (kind, level) slot: template type, and the level's
    grid for the stored levels; level 0's legs carry `_stream` when the
    streaming form runs them."""
    t = "float" if precision == "mixed" else "double"
    kind, level = entry["kind"], entry["level"]
    if kind == "k_fold":
        return f"k_fold[{('cycle', 'alpha', 'update', 'prepare', 'init', 'check')[level]}]"
    if kind in STEP_KERNELS or kind == "k_jpcg_small":
        return kind
    if kind in ("k_fine_down", "k_fine_up"):
        stream = plan and plan[0]["mode"] == "stream"
        return f"{kind}{'_stream' if stream else ''}<{t}>"
    if kind == "k_fine_dir":
        return f"{kind}<{t}>"
    grid = f"{plan[level]['nx']}x{plan[level]['ny']}" if level < len(plan) else "?"
    return f"{kind}<{t}> L{level} {grid}"


def load_traffic() -> dict:
    """Measured HBM bytes per launch by kernel, from the fenced JSON block of
    the committed counter summary ({} when absent)."""
    if not TRAFFIC_FILE.exists():
        return {}
    text = TRAFFIC_FILE.read_text()
    start = text.find("```json")
    if start < 0:
        return {}
    end = text.find("```", start + 7)
    try:
        return json.loads(text[start + 7:end])
    except ValueError:
        return {}


def run_path(precision: str, steps: int, warmup: int, nx: int, ny: int, tol: float,
             traffic: dict, tune_sweeps: bool = True, warm_start: int = 0):
    import numpy as np
    import torch

    from pycfd_amd import cases
    from pycfd_amd.analysis.postprocess import force_coefficients
    from pycfd_amd.physics.incompressible import Simulation

    cfg, body, u0, v0 = cases.cylinder(nx=nx, ny=ny, pressure_solver="mgcg",
                                       poisson_tol=tol, poisson_maxiter=300,
                                       precision=precision)
    t0 = time.perf_counter()
    sim = Simulation(cfg, obstacle=body, u_init=u0, v_init=v0)
    sim.solver.warm_start_order = warm_start
    log(f"[{precision}] setup {time.perf_counter() - t0:.1f} s "
        f"({sim.solver.pressure_solver.hierarchy})")
    plan = sim.solver.ctx.mg_plan()
    log(f"[{precision}] V-cycle plan: " + ", ".join(
        f"{p['nx']}x{p['ny']}:" + (f"fused {p['tile'][0]}x{p['tile'][1]}" if p["fused"]
                                   else p["mode"])
        for p in plan))
    if warmup:
        sim.run(max_steps=warmup)
    sweeps = {}
    if tune_sweeps:
        # the in-project tuner: multigrid smoothing sweeps timed on whole
        # steps from the warmed-up state, the fastest kept
        sweeps = sim.solver.tune_smoother_sweeps(sim.fields, sim.stepper.compute_dt(sim.fields),
                                                 candidates=(1, 2, 3), steps=3)
        log(f"[{precision}] smoother sweeps tuned: " + ", ".join(
            f"V({k},{k}) {v['ms_per_step']:.3f} ms/step {v['iterations']} it"
            for k, v in sweeps.items()) +
            f" -> V({sim.solver.pressure_solver.smooth_sweeps},"
            f"{sim.solver.pressure_solver.smooth_sweeps})")
        plan = sim.solver.ctx.mg_plan()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = sim.run(max_steps=steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    hist = res.history
    max_div = float(max(hist["max_div"]))
    cd_steps = np.array([force_coefficients((fx, fy), cfg.u_ref, cfg.l_ref)[0]
                         for fx, fy in zip(hist["fx"], hist["fy"])])
    out = {
        "mlups": nx * ny * steps / wall / 1e6,
        "ms_per_step": 1e3 * wall / steps,
        "max_div": max_div,
        "div_ok": max_div <= DIV_TARGET,
        "iterations_last_solve": int(sim.solver.pressure_solver.last_iterations),
        "cd_last": float(sim.force_coefficients()[0]),
        # mean drag coefficient over the timed steps, from each step's reaction
        "cd_mean": float(np.mean(cd_steps)),
        "mg_plan": [[p["nx"], p["ny"], p["mode"], *p["tile"]] for p in plan],
        "smoother_sweeps": int(sim.solver.pressure_solver.smooth_sweeps),
        "sweep_tuning_ms_per_step": {str(k): round(v["ms_per_step"], 3)
                                     for k, v in sweeps.items()},
        "t": float(sim.fields.t),
        "_cd_steps": cd_steps,  # per-step Cd, compared across paths, not printed
    }
    log(f"[{precision}] {steps} steps in {wall:.3f} s -> {out['mlups']:.1f} MLUPS, "
        f"max|div| {out['max_div']:.2e}, {out['iterations_last_solve']} it/solve")
    # second window, live kernel clock on
    ctx = sim.solver.ctx
    csteps = min(steps, 10)  # ~32 launches per iteration, ~16 iterations per solve
    ctx.kernel_clock_start(max_launches=csteps * 3 * 16 * 32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(max_steps=csteps)
    torch.cuda.synchronize()
    clocked = time.perf_counter() - t0
    kernels = []
    outside_us = 0.0  # kernel time per step outside the solve iterations
    for e in ctx.kernel_clock_read():
        name, n, ms, nbytes = kernel_name(e, precision, plan), e["launches"], e["ms"], e["bytes"]
        if e["kind"] in STEP_KERNELS[1:] or (e["kind"] == "k_fold" and e["level"] >= 3):
            outside_us += n * ms * 1e3 / csteps
        # walls, boundary passes and folds touch O(nx + ny) values or
        # partial sums: latency-bound, no bandwidth fraction
        achieved = nbytes / (ms * 1e-3) / 1e9 if nbytes else None
        kernels.append({"kernel": name, "bound": "hbm" if nbytes else "latency",
                        "achieved": round(achieved, 1) if nbytes else None,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4) if nbytes else None,
                        "traffic": traffic.get(name), "avg_us": round(ms * 1e3, 3),
                        "launches": n, "bytes_per_launch": nbytes,
                        "share_of_window": round(n * ms * 1e-3 / clocked, 4)})
    kernels.sort(key=lambda k: -k["share_of_window"])
    out["clocked_ms_per_step"] = 1e3 * clocked / csteps
    out["clocked_kernel_share"] = round(sum(k["share_of_window"] for k in kernels), 4)
    out["outside_solve_us_per_step"] = round(outside_us, 1)
    log(f"[{precision}] clocked window {1e3 * clocked / csteps:.3f} ms/step, clocked kernels "
        f"{100 * out['clocked_kernel_share']:.1f}% of it, {outside_us:.0f} us/step outside the "
        "solve iterations; top: " + "; ".join(
            f"{k['kernel']} {k['avg_us']} us x{k['launches']} frac {k['frac']} "
            f"share {k['share_of_window']}" for k in kernels[:8]))
    return sim, out, kernels


def other_workloads(streams: int, batch_steps: int, warm_start: int = 0) -> dict:
    """The other configured workloads, each a short timed window after a
    warm-up (not the headline): the 128 x 128 CG cavity, Taylor-Green 1024^2
    with multigrid V(2,2), and the 96 batched cavities advanced together.
    Each runs from the reference's start (the last pressure) and, when
    ``warm_start`` > 0, again with that extrapolation order (keys suffixed
    ``_warm``)."""
    import torch

    from pycfd_amd import cases
    from pycfd_amd.batch import SimulationBatch, cavity_batch_configs
    from pycfd_amd.physics.incompressible import Simulation
    out = {}

    def timed(sim, steps, warm):
        sim.run(max_steps=warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.run(max_steps=steps)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    orders = [0] + ([warm_start] if warm_start else [])
    for order in orders:
        sfx = "_warm" if order else ""
        cfg, _, _, _ = cases.cavity(nx=128, pressure_solver="cg", cfl_max=0.5, dt=0.01,
                                    poisson_tol=1e-10)
        sim = Simulation(cfg)
        sim.solver.warm_start_order = order
        w = timed(sim, 50, 5)
        key = "cavity128_cg" + sfx
        out[key] = {"mlups": round(128 * 128 * 50 / w / 1e6, 3),
                    "ms_per_step": round(1e3 * w / 50, 3),
                    "cg_iterations_last_solve": sim.solver.pressure_solver.last_iterations,
                    "warm_start_order": order}
        log(f"[{key}] {out[key]}")
        cfg, _, u0, v0 = cases.taylor_green(n=1024, nu=0.01, dt=1e-3,
                                            pressure_solver="multigrid", mg_sweeps=2)
        sim = Simulation(cfg, u_init=u0, v_init=v0)
        sim.solver.warm_start_order = order
        w = timed(sim, 40, 5)
        key = "taylor_green1024_mg_v22" + sfx
        out[key] = {
            "mlups": round(1024 * 1024 * 40 / w / 1e6, 2), "ms_per_step": round(1e3 * w / 40, 3),
            "cycles_last_solve": sim.solver.pressure_solver.last_iterations,
            "warm_start_order": order}
        log(f"[{key}] {out[key]}")
        del sim
        cfgs = cavity_batch_configs(96, seed=2024)
        batch = SimulationBatch(cfgs, streams=streams, warm_start_order=order)
        batch.run(max_steps=2)
        r = batch.run(max_steps=batch_steps)
        key = "batched_cavities96" + sfx
        out[key] = {
            "mlups": round(r.mlups, 2), "instances": len(cfgs), "streams": batch.n_streams,
            "steps": batch_steps, "wall_s": round(r.wall_time, 3),
            "setup_s": round(batch.setup_time, 2), "cells": int(sum(r.cells)),
            "sizes": sorted({c.nx for c in cfgs}), "seed": 2024, "warm_start_order": order}
        log(f"[{key}] {out[key]}")
        del batch
    return out


def copy_calibration(nbytes: int = 1 << 26) -> None:
    """A plain device copy of known size (nbytes read + nbytes written), run
    once so that a counter profile of this script holds a reference kernel
    for scaling FETCH_SIZE / WRITE_SIZE to HBM bytes."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()


def cpu_baseline(nx: int, ny: int, steps: int, tol: float):
    """The CPU oracle (NumPy/SciPy FP64 restatement of the reference) on the
    same workload, timed over `steps` steps after its own setup."""
    import numpy as np

    from oracle import oracle as O

    Lx, Ly = 32.0, 16.0
    dx, dy = Lx / nx, Ly / ny
    offs = (np.arange(8) + 0.5) / 8
    xf = np.concatenate(([0.0], np.cumsum(np.full(nx, dx))))
    yf = np.concatenate(([0.0], np.cumsum(np.full(ny, dy))))
    xf[-1], yf[-1] = Lx, Ly
    xs = (xf[:-1, None] + offs[None, :] * dx).ravel()
    ys = (yf[:-1, None] + offs[None, :] * dy).ravel()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    solid = (((X - 4.0) ** 2 + (Y - 8.0) ** 2 <= 0.25).astype(float)
             .reshape(nx, 8, ny, 8).mean(axis=(1, 3)) > 0.5)
    case = O.Case(nx=nx, ny=ny, lx=Lx, ly=Ly, re=100.0, dt=0.02, cfl_max=0.4,
                  kinds=("inlet", "outlet", "symmetry", "symmetry"),
                  velocity=(1.0, 0.0, 0.0, 0.0), solid=solid, pressure_solver="mgcg",
                  tol=tol)
    Xv, Yv = np.meshgrid(case.xc, case.yf, indexing="ij")
    v0 = 1e-3 * np.exp(-(((Xv - 5.0) / 1.0) ** 2 + ((Yv - 8.0) / 0.5) ** 2))
    st = O.State()
    u, v, p = O.initialize(case, st, 1.0, v0)
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = O.compute_dt(case, u, v)
        u, v, p = O.step(case, st, u, v, dt)
    wall = time.perf_counter() - t0
    return {"value": round(nx * ny * steps / wall / 1e6, 4), "unit": "MLUPS", "cores": 1,
            "kind": "port",
            "sample": f"{steps} RK3 steps of the {nx}x{ny} cylinder (mgcg, tol {tol:g}) after "
                      f"setup and the initial projection; {wall:.1f} s; NumPy/SciPy FP64, "
                      "single-threaded sparse kernels"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--paths", default="fp64,mixed")
    ap.add_argument("--tol", type=float, default=2e-11)
    ap.add_argument("--no-tune-sweeps", action="store_true",
                    help="keep V(1,1) (mg_sweeps=1) instead of the tuned sweeps")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the other workloads (cavity 128 CG, Taylor-Green 1024, batch)")
    ap.add_argument("--warm-start", type=int, default=4,
                    help="start each substep's solve from the extrapolation in time of its "
                         "last solutions (order 1-4; 0: the last pressure, the reference's)")
    ap.add_argument("--batch-streams", type=int, default=8)
    ap.add_argument("--batch-steps", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested); run with --gpus 1", file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 2
    torch.cuda.set_device(0)
    from pycfd_amd import _lib
    _lib.load()

    copy_calibration()
    traffic = load_traffic()
    results, kernels = {}, {}
    for prec in args.paths.split(","):
        _, results[prec], kernels[prec] = run_path(prec, args.steps, args.warmup, args.nx,
                                                   args.ny, args.tol, traffic,
                                                   not args.no_tune_sweeps, args.warm_start)
    main_path = "mixed" if "mixed" in results else "fp64"
    r = results[main_path]
    # the dominant kernel of the reported path: largest share of its window
    # among the bandwidth-bound ones
    dominant = dict(max((k for k in kernels[main_path] if k["bound"] == "hbm"),
                        key=lambda d: d["share_of_window"]))
    # traffic is not measured in this run: it is read from the committed
    # counter summary (two rocprofv3 --pmc passes, scaled by the copy kernel)
    dominant["traffic_source"] = (f"{TRAFFIC_FILE.relative_to(ROOT)} (rocprofv3 --pmc "
                                  "FETCH_SIZE / WRITE_SIZE per dispatch, scaled by a copy "
                                  "kernel of known size)" if dominant.get("traffic") else None)
    others = None if args.no_others else other_workloads(args.batch_streams, args.batch_steps,
                                                         args.warm_start)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(args.nx, args.ny, args.cpu_steps, args.tol)
        log(f"cpu baseline {cpu['value']} MLUPS")
    line = {
        "metric": "MLUPS per projection step @2048x1024, mixed vs FP64 at fixed div-u; HBM GB/s",
        "value": round(r["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_path == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"cylinder Re=100 {args.nx}x{args.ny} (32x16 D), OUTLET, "
                               f"mgcg to max|div u| <= {DIV_TARGET:g} (tol {args.tol:g}), "
                               "SSP-RK3",
                   "path": main_path, "nx": args.nx, "ny": args.ny,
                   "warm_start": (f"order-{args.warm_start} extrapolation in time per substep"
                                  if args.warm_start else "previous pressure")},
        "paths": {k: {kk: (float(f"{vv:.5g}") if isinstance(vv, float) else vv)
                      for kk, vv in v.items() if not kk.startswith("_")}
                  for k, v in results.items()},
        "roofline": dominant,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "other_workloads": others,
    }
    if "fp64" in results and "mixed" in results:
        line["mixed_speedup_vs_fp64"] = round(results["mixed"]["mlups"] / results["fp64"]["mlups"], 3)
        # the same steps on both paths: mean Cd of the mixed path against
        # FP64, and the largest per-step relative difference
        cf, cm = results["fp64"]["cd_mean"], results["mixed"]["cd_mean"]
        import numpy as np
        sf, sm = results["fp64"]["_cd_steps"], results["mixed"]["_cd_steps"]
        if cf and len(sf) == len(sm):
            line["cd_mean_rel_diff_mixed_vs_fp64"] = float(f"{abs(cm - cf) / abs(cf):.3g}")
            line["cd_step_max_rel_diff_mixed_vs_fp64"] = float(
                f"{float(np.max(np.abs(sm - sf) / np.abs(sf))):.3g}")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
