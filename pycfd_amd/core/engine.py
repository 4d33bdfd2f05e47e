"""One device context of ``libpycfd_hip.so``: the handle a ProjectionSolver
(or a standalone pressure solver) drives.  Thin ctypes plumbing only; the
arithmetic all happens in the HIP kernels."""

from __future__ import annotations

import atexit
import ctypes as C
import sys
import weakref

import numpy as np

from .. import _lib

# Contexts still open.  Each holds device memory, events and pinned host
# memory of the HIP runtime, so each is destroyed while the runtime is
# certainly alive: by close(), by garbage collection during the program, or
# by the exit hook below -- never from a finaliser that runs after the
# interpreter (and with it torch's and the runtime's own exit handlers) has
# begun to shut down, where a HIP call can touch a finalised runtime.
_LIVE: "weakref.WeakSet[DeviceContext]" = weakref.WeakSet()
_EXIT_HOOKED = False


def _close_all() -> None:
    for ctx in list(_LIVE):
        ctx.close()


def _track(ctx: "DeviceContext") -> None:
    """Register the exit hook on first use: after torch initialised the GPU,
    so (exit hooks run last-registered first) it runs before torch's."""
    global _EXIT_HOOKED
    if not _EXIT_HOOKED:
        atexit.register(_close_all)
        _EXIT_HOOKED = True
    _LIVE.add(ctx)


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(0 if t is None else int(t.data_ptr()))


def require_device() -> None:
    """Fail loudly unless a HIP device is usable through torch."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError(
            "pycfd_amd needs a HIP device: torch.cuda.is_available() is False "
            "(the HIP engine has no CPU fallback)"
        )


class DeviceContext:
    """``pcfd_ctx`` plus the arrays it was created from (kept alive here)."""

    def __init__(self, nx: int, ny: int, dx: float, dy: float, nu: float,
                 body_force=(0.0, 0.0), scheme: str = "rk3",
                 bc_kinds=("no_slip",) * 4, bc_velocity=(0.0,) * 4, bc_p_ref=(0.0,) * 4,
                 inlet_lines=(None,) * 4, solid: np.ndarray | None = None,
                 body_id: np.ndarray | None = None, n_bodies: int = 0,
                 les_coeff: float | None = None, *, les_cs: float = 0.0,
                 dx_cells: np.ndarray | None = None, dy_cells: np.ndarray | None = None,
                 area_mean: float | None = None, soft_ib: bool = False) -> None:
        require_device()
        import torch
        self._lib = _lib.load()
        self.nx, self.ny = int(nx), int(ny)
        d = _lib.CaseDesc()
        d.nx, d.ny = self.nx, self.ny
        d.dx, d.dy, d.nu = float(dx), float(dy), float(nu)
        d.body_fx, d.body_fy = float(body_force[0]), float(body_force[1])
        d.time_scheme = _lib.SCHEME_CODES[scheme]
        self._keep = []
        for w in range(4):
            d.bc_kind[w] = _lib.BC_CODES[bc_kinds[w]]
            d.bc_velocity[w] = float(bc_velocity[w])
            d.bc_p_ref[w] = float(bc_p_ref[w])
            line = inlet_lines[w]
            if line is not None:
                arr = np.ascontiguousarray(line, dtype=np.float64)
                self._keep.append(arr)
                d.inlet_line[w] = arr.ctypes.data_as(C.POINTER(C.c_double))
        if solid is not None and np.any(solid):
            sarr = np.ascontiguousarray(solid, dtype=np.uint8)
            self._keep.append(sarr)
            d.solid = sarr.ctypes.data_as(C.POINTER(C.c_uint8))
        if body_id is not None:
            barr = np.ascontiguousarray(body_id, dtype=np.int8)
            self._keep.append(barr)
            d.body_id = barr.ctypes.data_as(C.POINTER(C.c_int8))
        d.n_bodies = int(n_bodies)
        self.n_bodies = int(n_bodies) if n_bodies > 1 else 0
        d.use_les = int(les_coeff is not None)
        d.les_coeff = float(les_coeff or 0.0)
        d.les_cs = float(les_cs)
        self.stretched = dx_cells is not None
        if self.stretched:
            if dy_cells is None or area_mean is None:
                raise ValueError("a stretched mesh needs dx_cells, dy_cells and area_mean")
            wx = np.ascontiguousarray(dx_cells, dtype=np.float64)
            wy = np.ascontiguousarray(dy_cells, dtype=np.float64)
            if wx.shape != (self.nx,) or wy.shape != (self.ny,):
                raise ValueError(f"cell widths must have shapes ({self.nx},), ({self.ny},)")
            self._keep += [wx, wy]
            d.dx_cells = wx.ctypes.data_as(C.POINTER(C.c_double))
            d.dy_cells = wy.ctypes.data_as(C.POINTER(C.c_double))
            d.area_mean = float(area_mean)
        d.soft_ib = int(bool(soft_ib))
        self.soft_ib = bool(soft_ib)
        self.stream = torch.cuda.current_stream()
        handle = C.c_void_p(0)
        _lib.check(self._lib.pcfd_ctx_create(C.byref(d), C.c_void_p(self.stream.cuda_stream),
                                             C.byref(handle)), "pcfd_ctx_create")
        self.handle = handle
        self._keep = []
        _track(self)

    def close(self) -> None:
        """Destroy the device context (waits for its stream first)."""
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.pcfd_ctx_destroy(self.handle)
            self.handle = None
        _LIVE.discard(self)

    def __del__(self):
        if sys.is_finalizing():
            return  # the exit hook closed every context the program left open
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name: str, *args) -> None:
        _lib.check(getattr(self._lib, name)(self.handle, *args), name)

    # -- configuration ---------------------------------------------------
    def set_solver(self, kind: str, tol: float, maxiter: int, omega: float, sweeps: int,
                   precision: int, levels: list[dict] | None, *, fused_levels: bool = True,
                   tile: tuple[int, int] = (0, 0), coarse_chain: int = 0,
                   stream_rows: int = 0, row_passes: int = 0) -> None:
        sd = _lib.SolverDesc(_lib.SOLVER_CODES[kind], float(tol), int(maxiter), float(omega),
                             int(sweeps), int(precision), int(bool(fused_levels)),
                             int(tile[0]), int(tile[1]), int(coarse_chain), int(stream_rows),
                             int(row_passes))
        keep = []
        if levels:
            arr = (_lib.MGLevelDesc * len(levels))()
            for i, L in enumerate(levels):
                e = arr[i]
                e.nx, e.ny, e.fx, e.fy = L["nx"], L["ny"], L["fx"], L["fy"]
                e.weight = float(L["weight"])
                act = np.ascontiguousarray(L["active"], dtype=np.uint8)
                keep.append(act)
                e.active = act.ctypes.data_as(C.POINTER(C.c_uint8))
                if L.get("table") is not None:
                    di = np.ascontiguousarray(L["di"], dtype=np.int8)
                    dj = np.ascontiguousarray(L["dj"], dtype=np.int8)
                    wide = np.asarray(L["cls"]).dtype == np.uint32
                    cls = np.ascontiguousarray(L["cls"], dtype=np.uint32 if wide else np.uint16)
                    tab = np.ascontiguousarray(L["table"], dtype=np.float64)
                    cdv = np.ascontiguousarray(L["class_dinv"], dtype=np.float64)
                    cac = np.ascontiguousarray(L["class_active"], dtype=np.uint8)
                    keep += [di, dj, cls, tab, cdv, cac]
                    e.n_offsets = di.size
                    e.di = di.ctypes.data_as(C.POINTER(C.c_int8))
                    e.dj = dj.ctypes.data_as(C.POINTER(C.c_int8))
                    e.n_classes = cdv.size
                    if wide:
                        e.cls32 = cls.ctypes.data_as(C.POINTER(C.c_uint32))
                    else:
                        e.cls = cls.ctypes.data_as(C.POINTER(C.c_uint16))
                    e.table = tab.ctypes.data_as(C.POINTER(C.c_double))
                    e.class_dinv = cdv.ctypes.data_as(C.POINTER(C.c_double))
                    e.class_active = cac.ctypes.data_as(C.POINTER(C.c_uint8))
                if L.get("coarse_inv") is not None:
                    ci = np.ascontiguousarray(L["coarse_inv"], dtype=np.float64)
                    keep.append(ci)
                    e.coarse_n = int(ci.shape[0])
                    e.coarse_inv = ci.ctypes.data_as(C.POINTER(C.c_double))
            self._call("pcfd_set_pressure_solver", C.byref(sd), arr, len(levels))
        else:
            self._call("pcfd_set_pressure_solver", C.byref(sd), None, 0)

    # -- operators -------------------------------------------------------
    def momentum_rhs(self, u, v, blend: float, rhs_u, rhs_v) -> None:
        self._call("pcfd_momentum_rhs", _ptr(u), _ptr(v), float(blend), _ptr(rhs_u),
                   _ptr(rhs_v))

    def eddy_viscosity(self, u, v, nu_c, nu_corner=None) -> None:
        self._call("pcfd_eddy_viscosity", _ptr(u), _ptr(v), _ptr(nu_c),
                   _ptr(nu_corner) if nu_corner is not None else None)

    def momentum_rhs_variable(self, u, v, blend: float, nu_c, rhs_u, rhs_v) -> None:
        self._call("pcfd_momentum_rhs_variable", _ptr(u), _ptr(v), float(blend), _ptr(nu_c),
                   _ptr(rhs_u), _ptr(rhs_v))

    def divergence(self, u, v, out) -> None:
        self._call("pcfd_divergence", _ptr(u), _ptr(v), _ptr(out))

    def apply_velocity_bc(self, u, v, predictor: bool) -> None:
        self._call("pcfd_apply_velocity_bc", _ptr(u), _ptr(v), int(bool(predictor)))

    def mass_balance(self, u, v) -> None:
        self._call("pcfd_mass_balance", _ptr(u), _ptr(v))

    def mask_obstacle(self, u, v, dt: float | None, mode: int) -> None:
        self._call("pcfd_mask_obstacle", _ptr(u), _ptr(v), float(dt or 0.0), int(mode))

    def wrap_faces(self, u, v) -> None:
        self._call("pcfd_wrap_faces", _ptr(u), _ptr(v))

    def apply_pressure_bc(self, p) -> None:
        self._call("pcfd_apply_pressure_bc", _ptr(p))

    def pressure_solve(self, rhs, p0, out) -> _lib.SolveInfo:
        info = _lib.SolveInfo()
        self._call("pcfd_pressure_solve", _ptr(rhs), _ptr(p0), _ptr(out), C.byref(info))
        return info

    def project(self, u, v, dt: float, p) -> _lib.SolveInfo:
        info = _lib.SolveInfo()
        self._call("pcfd_project", _ptr(u), _ptr(v), float(dt), _ptr(p), C.byref(info))
        return info

    def initialize(self, u, v, p) -> _lib.SolveInfo:
        info = _lib.SolveInfo()
        self._call("pcfd_initialize", _ptr(u), _ptr(v), _ptr(p), C.byref(info))
        return info

    def step(self, u0, v0, dt: float, blend: float, u, v, p, want_info: bool = True):
        info = _lib.SolveInfo() if want_info else None
        self._call("pcfd_step", _ptr(u0), _ptr(v0), float(dt), float(blend), _ptr(u), _ptr(v),
                   _ptr(p), C.byref(info) if info is not None else None)
        return info

    def reset_guess(self) -> None:
        self._call("pcfd_reset_guess")

    def set_warm_start(self, order: int) -> None:
        """Warm start of each substep's solve by extrapolation in time of
        that substep's last solutions (order 1-3; 0: the previous pressure,
        the reference's) (pcfd_set_warm_start)."""
        self._call("pcfd_set_warm_start", int(order))

    def set_step_graphs(self, enable: bool) -> None:
        """Capture whole steps as HIP graphs and replay them (pcfd_set_step_graphs)."""
        self._call("pcfd_set_step_graphs", int(bool(enable)))

    def step_graph_stats(self) -> dict:
        """Graph replays, captures and eager redos of this context so far."""
        out = (C.c_int32 * 3)()
        self._call("pcfd_step_graph_stats", out)
        return {"launches": int(out[0]), "captures": int(out[1]), "fallbacks": int(out[2])}

    def pressure_guess(self, buf, store: bool) -> None:
        """Copy the warm-start pressure out to (store=False) or in from ``buf``."""
        self._call("pcfd_pressure_guess", _ptr(buf), int(bool(store)))

    def diagnostics(self, u, v, p, u_prev=None, v_prev=None) -> np.ndarray:
        out = np.zeros(16)
        self._call("pcfd_diagnostics", _ptr(u), _ptr(v), _ptr(p), _ptr(u_prev), _ptr(v_prev),
                   out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def force(self) -> np.ndarray:
        out = np.zeros(2 + 2 * self.n_bodies)
        self._call("pcfd_get_force", out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def poisson_apply(self, x, y) -> None:
        self._call("pcfd_poisson_apply", _ptr(x), _ptr(y))

    def mg_apply(self, r, z) -> None:
        self._call("pcfd_mg_apply", _ptr(r), _ptr(z))

    def time_kernel(self, which: int, reps: int) -> tuple[float, float]:
        """(mean ms per launch, algorithmic bytes per launch) of one hot kernel."""
        ms, b = C.c_double(0), C.c_double(0)
        self._call("pcfd_time_kernel", int(which), int(reps), C.byref(ms), C.byref(b))
        return float(ms.value), float(b.value)

    CLOCK_KINDS = ("k_cg_update", "k_fine_dir", "k_fine_down", "k_fine_up", "k_class_down",
                   "k_class_up", "k_cls_rows[smooth0+res]", "k_cls_rows[smooth]",
                   "k_cls_rows[res]", "k_cls_rows[prolong]", "k_cls_restrict", "k_dense_apply",
                   "k_chain", "k_fold", "k_predictor", "k_predictor_comb", "k_walls",
                   "k_boundary_pass", "k_prepare", "k_init", "k_check", "k_scatter",
                   "k_correct_ghosts", "k_combine", "k_diagnostics", "k_jpcg_small",
                   "k_warm_guess")
    CLOCK_LEVELS = 16

    def kernel_clock_start(self, max_launches: int) -> None:
        """Time every kernel of the steps and pressure solves that follow, live."""
        self._call("pcfd_kernel_clock_start", int(max_launches))

    def kernel_clock_read(self) -> list[dict]:
        """Per clocked (kind, level): launches, mean ms, algorithmic bytes per launch."""
        slots = len(self.CLOCK_KINDS) * self.CLOCK_LEVELS
        n = (C.c_int32 * slots)()
        ms, b = (C.c_double * slots)(), (C.c_double * slots)()
        self._call("pcfd_kernel_clock_read", n, ms, b)
        out = []
        for q in range(slots):
            if n[q]:
                kind, level = divmod(q, self.CLOCK_LEVELS)
                out.append({"kind": self.CLOCK_KINDS[kind], "level": level, "launches": int(n[q]),
                            "ms": float(ms[q]), "bytes": float(b[q])})
        return out

    def mg_plan(self) -> list[dict]:
        """Per multigrid level: grid, whether its legs run fused, and the tile."""
        buf = (C.c_int32 * (5 * 32))()
        n = int(self._lib.pcfd_mg_plan(self.handle, buf, len(buf)))
        if n < 0:
            _lib.check(-n, "pcfd_mg_plan")
        vals = list(buf)
        modes = {0: "passes", 1: "fused", 2: "chain", 3: "stream"}
        return [{"nx": vals[5 * k], "ny": vals[5 * k + 1], "mode": modes[vals[5 * k + 2]],
                 "fused": vals[5 * k + 2] in (1, 3),
                 "tile": (vals[5 * k + 3], vals[5 * k + 4])}
                for k in range(min(n, 32))]

    def synchronize(self) -> None:
        self._call("pcfd_synchronize")
