"""The small cases the golden vectors were recorded on, in both vocabularies:
oracle ``Case`` keyword arguments and product ``SimulationConfig`` builders."""

from __future__ import annotations

import numpy as np

from oracle.oracle import Case

PERIODIC = ("periodic",) * 4

CASES = {
    "cavity": dict(nx=24, ny=20, lx=1.0, ly=1.0, re=100.0, dt=0.01, cfl_max=0.5,
                   kinds=("no_slip", "no_slip", "no_slip", "moving_wall"),
                   velocity=(0.0, 0.0, 0.0, 1.0)),
    "periodic": dict(nx=24, ny=16, lx=2 * np.pi, ly=2 * np.pi, re=50.0, dt=2e-3,
                     adaptive_dt=False, kinds=PERIODIC),
    "channel": dict(nx=32, ny=16, lx=4.0, ly=1.0, re=50.0, dt=0.01, cfl_max=0.4,
                    kinds=("inlet", "outlet", "no_slip", "no_slip"),
                    velocity=(1.0, 0.0, 0.0, 0.0),
                    profile=("parabolic", "uniform", "uniform", "uniform")),
    "cyl_outlet": dict(nx=48, ny=24, lx=12.0, ly=6.0, re=100.0, dt=0.02, cfl_max=0.4,
                       kinds=("inlet", "outlet", "symmetry", "symmetry"),
                       velocity=(1.0, 0.0, 0.0, 0.0)),
    "cyl_pout": dict(nx=48, ny=24, lx=12.0, ly=6.0, re=100.0, dt=0.02, cfl_max=0.4,
                     kinds=("inlet", "pressure_outlet", "symmetry", "symmetry"),
                     velocity=(1.0, 0.0, 0.0, 0.0)),
}

# Cases compared with the oracle only (the reference recorded nothing for them).
ORACLE_CASES = {
    # body-force channel, periodic in x (channel_flow.py periodic mode)
    "pchannel": dict(nx=16, ny=24, lx=2.0, ly=1.0, re=10.0, dt=0.01, cfl_max=0.4,
                     kinds=("periodic", "periodic", "no_slip", "no_slip"),
                     body_force=(0.8, 0.0)),
}


def _base(base):
    return CASES[base] if base in CASES else ORACLE_CASES[base]


# Stretched variants of the cases above: base case plus stretching
# (tests/golden/make_golden.py records the same four).
STRETCHED = {
    "s_cavity": dict(base="cavity", stretch_x=1.06, stretch_y=1.04, cluster_x="walls",
                     cluster_y="walls"),
    "s_channel": dict(base="channel", stretch_y=1.08, cluster_y="walls"),
    "s_cyl": dict(base="cyl_outlet", stretch_x=1.03, stretch_y=1.05, cluster_x="low",
                  cluster_y="centre"),
    "s_cylp": dict(base="cyl_pout", stretch_y=1.05, cluster_y="centre"),
}

# Stretched cases checked against the oracle only (no reference recording):
# a uniform periodic axis beside a stretched one.
STRETCHED_ORACLE = {
    "s_pchannel": dict(base="pchannel", stretch_y=1.08, cluster_y="walls"),
}


def _split(name):
    for table in (STRETCHED, STRETCHED_ORACLE):
        if name in table:
            spec = dict(table[name])
            return spec.pop("base"), spec
    return name, {}


def circle(nx, ny, lx, ly, cx, cy, r, sub=8, dx_cells=None, dy_cells=None):
    """Solid iff more than half of 8x8 mid-point samples lie in the circle
    (obstacles.py:87-135)."""
    offs = (np.arange(sub) + 0.5) / sub
    dxc = np.full(nx, lx / nx) if dx_cells is None else np.asarray(dx_cells)
    dyc = np.full(ny, ly / ny) if dy_cells is None else np.asarray(dy_cells)
    xf = np.concatenate(([0.0], np.cumsum(dxc)))
    yf = np.concatenate(([0.0], np.cumsum(dyc)))
    xf[-1], yf[-1] = lx, ly
    xs = (xf[:-1, None] + offs[None, :] * dxc[:, None]).ravel()
    ys = (yf[:-1, None] + offs[None, :] * dyc[:, None]).ravel()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    hits = ((X - cx) ** 2 + (Y - cy) ** 2 <= r ** 2).astype(float)
    return hits.reshape(nx, sub, ny, sub).mean(axis=(1, 3)) > 0.5


def oracle_case(name: str, **kw) -> Case:
    base, stretch = _split(name)
    args = dict(_base(base))
    args.update(stretch)
    args.update(kw)
    case = Case(**args)
    if base.startswith("cyl"):
        case = Case(**args, solid=circle(args["nx"], args["ny"], args["lx"], args["ly"], 3.0,
                                         3.0, 0.5, dx_cells=case.dx_cells,
                                         dy_cells=case.dy_cells))
    return case


def initial(name: str, case: Case):
    """Initial faces used by make_golden.py for the trajectory runs."""
    name = _split(name)[0]
    if name.startswith("cyl") or name == "channel":
        Xv, Yv = np.meshgrid(case.xc, case.yf, indexing="ij")
        v0 = 0.05 * np.exp(-((Xv - 4.0) ** 2 + (Yv - 3.0) ** 2))
        return (1.0 if name.startswith("cyl") else 0.0), v0
    if name == "periodic":
        Xu, Yu = np.meshgrid(case.xf, case.yc, indexing="ij")
        Xv, Yv = np.meshgrid(case.xc, case.yf, indexing="ij")
        return np.sin(Xu) * np.cos(Yu), -np.cos(Xv) * np.sin(Yv)
    if name == "pchannel":
        Xv, Yv = np.meshgrid(case.xc, case.yf, indexing="ij")
        return None, 0.05 * np.sin(np.pi * Xv) * np.sin(np.pi * Yv)
    return None, None


def product_config(name: str, **kw):
    """SimulationConfig + obstacle for the same case, built with pycfd_amd."""
    from pycfd_amd.config import BCKind, BCSpec, SimulationConfig
    from pycfd_amd.core.mesh import StructuredMesh
    from pycfd_amd.geometry.obstacles import circle_mask
    base, stretch = _split(name)
    c = _base(base)
    walls = ("left", "right", "bottom", "top")
    prof = c.get("profile", ("uniform",) * 4)
    vel = c.get("velocity", (0.0,) * 4)
    bcs = {w: BCSpec(BCKind(c["kinds"][k]), velocity=vel[k], profile=prof[k])
           for k, w in enumerate(walls)}
    args = dict(nx=c["nx"], ny=c["ny"], lx=c["lx"], ly=c["ly"], re=c["re"], dt=c["dt"],
                cfl_max=c.get("cfl_max", 0.5), adaptive_dt=c.get("adaptive_dt", True),
                t_end=100.0, boundary_config=bcs, name=name,
                body_force=tuple(c.get("body_force", (0.0, 0.0))))
    args.update(stretch)
    args.update(kw)
    cfg = SimulationConfig(**args)
    body = None
    if base.startswith("cyl"):
        body = circle_mask(StructuredMesh.from_config(cfg), (3.0, 3.0), 0.5)
    return cfg, body
