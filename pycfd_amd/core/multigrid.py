"""Aggregation-multigrid hierarchy: native host build, device V-cycle.

The hierarchy of an operator is built once by the library's native host
builder (``pcfd_mg_build``, csrc/hierarchy.hip) with the reference's algorithm
(``pycfd/core/multigrid.py:324-474``): geometric 2x2 / 2x1 / 1x2 aggregates
that keep each level a rectangular grid with holes (:167-245), the 0/1
tentative prolongator smoothed by one damped Jacobi pass (:248-273),
Galerkin coarse operators with the identity on dropped aggregates (:276-287,
380-383), damping weights from a seeded 12-step power iteration (:129-157,
354-362), and the coarsest level solved exactly as a pinned solve followed
by mean removal (:514-543).

Every level is a structured grid, so the builder forms each distinct coarse
row once (rows whose neighbourhoods carry the same operator rows are equal up
to a translation) and hands each coarse operator to the GPU as a row
dictionary: a per-cell class index into the level's distinct operator rows.
The prolongator is never stored: the device applies ``P = (I - w D^-1 A)
P_tent`` and ``P^T`` on the fly from the level's own operator.  The device's
coarsest level is the first one with at most ``coarse_max`` unknowns, solved
by its dense FP64 pseudo-inverse; with ``coarse_max`` at the reference's 96
this is the reference's own coarsest level.  A hierarchy is built only as deep
as that level.
"""

from __future__ import annotations

import ctypes as C
import hashlib
import logging
import os
import threading
import weakref

import numpy as np
import scipy.sparse as sp

from .. import _lib

log = logging.getLogger(__name__)

COARSEST_UNKNOWNS = 96

# Host threads of the native build (the result does not depend on them).
_SETUP_THREADS = max(1, min(16, os.cpu_count() or 1))

_START_LOCK = threading.Lock()
_START: np.ndarray = np.zeros(0)


def _start_vector(n: int) -> np.ndarray:
    """``default_rng(0).standard_normal(n)``, the power iteration's start
    (multigrid.py:147-149).  A generator draws its normals in sequence, so
    every shorter start vector is a prefix of a longer one: one vector, grown
    on demand, serves every level and operator."""
    global _START
    with _START_LOCK:
        if _START.size < n:
            _START = np.random.default_rng(0).standard_normal(n)
        return _START


class _Built:
    """The native hierarchy (``pcfd_mg_setup``), freed with this object."""

    def __init__(self, handle: C.c_void_p) -> None:
        self.handle = handle
        self._free = _lib.load().pcfd_mg_free  # still reachable at interpreter exit

    def __del__(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._free(self.handle)
            self.handle = None

    def info(self, level: int) -> _lib.MGBuiltLevel:
        out = _lib.MGBuiltLevel()
        _lib.check(_lib.load().pcfd_mg_level_info(self.handle, level, C.byref(out)),
                   "pcfd_mg_level_info")
        return out

    def copy(self, level: int, *fields: str) -> dict[str, np.ndarray]:
        """The named arrays of one level (pcfd_mg_level_copy)."""
        inf = self.info(level)
        n = inf.nx * inf.ny
        shapes = {
            "indptr": (n + 1, np.int64), "indices": (inf.nnz, np.int32),
            "data": (inf.nnz, np.float64), "dinv": (n, np.float64), "active": (n, np.uint8),
            "di": (inf.n_offsets, np.int8), "dj": (inf.n_offsets, np.int8),
            "cls": (n if inf.n_classes else 0, np.uint32),
            "table": (inf.n_classes * inf.n_offsets, np.float64),
            "class_dinv": (inf.n_classes, np.float64), "class_active": (inf.n_classes, np.uint8),
        }
        out = {f: np.empty(shapes[f][0], dtype=shapes[f][1]) for f in fields}
        ptrs = [C.c_void_p(out[f].ctypes.data) if f in out and out[f].size else None
                for f in shapes]
        _lib.check(_lib.load().pcfd_mg_level_copy(self.handle, level, *ptrs),
                   "pcfd_mg_level_copy")
        return out


class Level:
    """One rung of the hierarchy; its arrays are fetched from the native
    build when first asked for."""

    def __init__(self, built: _Built, index: int, info: _lib.MGBuiltLevel,
                 singular: bool) -> None:
        self._built = built
        self.index = index
        self.shape = (int(info.nx), int(info.ny))
        self.factors = (int(info.fx), int(info.fy))
        self.weight = float(info.weight)
        self.nnz = int(info.nnz)
        self.pin = int(info.pin)
        self.singular = singular
        self._cache: dict[str, object] = {}

    @property
    def n(self) -> int:
        return self.shape[0] * self.shape[1]

    def _get(self, name: str):
        if name not in self._cache:
            self._cache.update(self._built.copy(self.index, name))
        return self._cache[name]

    @property
    def A(self) -> sp.csr_matrix:
        """The level's operator (columns ascending, no stored zeros)."""
        if "A" not in self._cache:
            d = self._built.copy(self.index, "indptr", "indices", "data")
            self._cache["A"] = sp.csr_matrix((d["data"], d["indices"], d["indptr"]),
                                             shape=(self.n, self.n))
        return self._cache["A"]

    @property
    def dinv(self) -> np.ndarray:
        return self._get("dinv")

    @property
    def active(self) -> np.ndarray:
        if "active_bool" not in self._cache:
            self._cache["active_bool"] = self._get("active").astype(bool)
        return self._cache["active_bool"]


def _inverse_negated(B: np.ndarray) -> np.ndarray:
    """inv(-B) for a symmetric positive definite B (overwritten), by Cholesky
    (half the work of a general inverse); a general inverse when the
    factorisation fails."""
    from scipy.linalg import lapack
    n = B.shape[0]
    if n == 0:
        return B
    # B^T is B's memory in column order, so LAPACK works on it in place
    c, info = lapack.dpotrf(B.T, lower=1, overwrite_a=1, clean=0)
    if info == 0:
        inv, info = lapack.dpotri(c, lower=1, overwrite_c=1)
        if info == 0:
            for i0 in range(0, n, 256):  # the lower triangle onto the upper
                i1 = min(n, i0 + 256)
                blk = inv[i0:i1, i0:i1]
                iu = np.triu_indices(i1 - i0, 1)
                blk[iu] = blk.T[iu]
                inv[i0:i1, i1:] = inv[i1:, i0:i1].T
            np.negative(inv, out=inv)
            return inv
    return np.linalg.inv(-B)


def coarsest_pseudo_inverse(level: Level) -> np.ndarray:
    """Dense operator of the coarsest solve, projections included (multigrid.py:514-543).

    Singular case: ``Pi inv(M) Z Pi`` with M the operator pinned at ``pin``,
    Z zeroing the pinned right-hand side entry and Pi the mean removal on the
    active set (zero on inactive cells); the two projections are applied as
    column and row mean subtractions.  Otherwise inactive rows are zeroed.

    Inactive cells carry identity rows that no active row couples to, and
    pinning cell p leaves inv(M) with column p of no effect once Z drops it,
    so only the inverse of the operator on the active cells (without p) is
    needed; that block is symmetric negative definite and is inverted by
    Cholesky.
    """
    n = level.n
    ia = np.flatnonzero(level.active)
    if ia.size == 0:
        return np.zeros((n, n))
    A = level.A
    m = ia.size
    if level.singular:
        p = int(np.searchsorted(ia, level.pin))  # the pin's place among the active cells
        keep = np.delete(ia, p)
        Z = _inverse_negated((-A[keep][:, keep]).toarray())
        Y = np.empty((m, m))  # Z with a zero row and column at p
        Y[p, :] = 0.0
        Y[:, p] = 0.0
        Y[:p, :p] = Z[:p, :p]
        Y[:p, p + 1:] = Z[:p, p:]
        Y[p + 1:, :p] = Z[p:, :p]
        Y[p + 1:, p + 1:] = Z[p:, p:]
        Y -= Y.mean(axis=0, keepdims=True)
        Y -= Y.mean(axis=1, keepdims=True)
    else:
        Y = _inverse_negated((-A[ia][:, ia]).toarray())
    if m == n:
        return np.ascontiguousarray(Y)
    out = np.zeros((n, n))
    out[np.ix_(ia, ia)] = Y
    return out


class MultigridHierarchy:
    """A V-cycle hierarchy for one operator (multigrid.py:324-474), built down
    to the first level with at most ``coarse_max`` unknowns (the reference's
    whole hierarchy at 96)."""

    def __init__(self, A: sp.csr_matrix, shape: tuple[int, int], active: np.ndarray,
                 singular: bool, presmooth: int = 1, postsmooth: int = 1,
                 periodic: tuple[bool, bool] = (False, False),
                 coarse_max: int = COARSEST_UNKNOWNS) -> None:
        if presmooth < 1 or postsmooth < 1:
            raise ValueError(
                "a V-cycle needs at least one smoothing sweep on each side, got "
                f"pre={presmooth}, post={postsmooth}")
        if presmooth != postsmooth:
            raise ValueError(
                f"a V-cycle with pre={presmooth} and post={postsmooth} sweeps is "
                "not a symmetric operator, and conjugate gradient needs one; use "
                "an equal number on each side")
        self.presmooth = presmooth
        self.postsmooth = postsmooth
        self.periodic = (bool(periodic[0]), bool(periodic[1]))
        self.singular = bool(singular)
        self._A = sp.csr_matrix(A)
        self._shape = (int(shape[0]), int(shape[1]))
        self._active = np.ascontiguousarray(np.asarray(active, dtype=bool).ravel(),
                                            dtype=np.uint8)
        self._lock = threading.Lock()
        self._device: dict[tuple[int, bool], list[dict]] = {}
        self._build(int(coarse_max))

    def _build(self, coarse_max: int) -> None:
        A = self._A
        nx, ny = self._shape
        n = nx * ny
        if A.shape != (n, n) or self._active.size != n:
            raise ValueError(f"operator {A.shape} and active mask ({self._active.size}) "
                             f"do not fit a {nx}x{ny} grid")
        indptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        indices = np.ascontiguousarray(A.indices, dtype=np.int32)
        data = np.ascontiguousarray(A.data, dtype=np.float64)
        x0 = _start_vector(n)
        op = _lib.MGOperator()
        op.nx, op.ny = nx, ny
        op.indptr = indptr.ctypes.data_as(C.POINTER(C.c_int64))
        op.indices = indices.ctypes.data_as(C.POINTER(C.c_int32))
        op.data = data.ctypes.data_as(C.POINTER(C.c_double))
        op.active = self._active.ctypes.data_as(C.POINTER(C.c_uint8))
        op.x0 = x0.ctypes.data_as(C.POINTER(C.c_double))
        op.singular = int(self.singular)
        op.periodic_x, op.periodic_y = int(self.periodic[0]), int(self.periodic[1])
        op.threads = _SETUP_THREADS
        op.coarse_max = coarse_max
        op.every_row = 0
        handle = C.c_void_p(0)
        _lib.check(_lib.load().pcfd_mg_build(C.byref(op), C.byref(handle)), "pcfd_mg_build")
        built = _Built(handle)
        first = built.info(0)
        self.levels = [Level(built, l, first if l == 0 else built.info(l), self.singular)
                       for l in range(first.n_levels)]
        self.levels[0]._cache["A"] = A  # the build keeps no copy of the fine operator
        self._built = built
        self._depth = coarse_max
        log.debug("multigrid hierarchy: %s", self)

    def _ensure_depth(self, coarse_max: int) -> None:
        """Rebuild deeper when a coarser device level than this build's is
        asked for (the levels above it come out the same)."""
        last = self.levels[-1]
        if coarse_max < self._depth and last.n > max(coarse_max, COARSEST_UNKNOWNS) \
                and last.n <= self._depth:
            self._build(coarse_max)

    @property
    def n_levels(self) -> int:
        return len(self.levels)

    @property
    def unknowns(self) -> list[int]:
        return [lv.n for lv in self.levels]

    @property
    def operator_complexity(self) -> float:
        fine = self.levels[0].nnz
        return sum(lv.nnz for lv in self.levels) / fine if fine else 1.0

    def __repr__(self) -> str:
        return (f"MultigridHierarchy({self.n_levels} levels, {self.unknowns[0]} -> "
                f"{self.unknowns[-1]} unknowns, complexity {self.operator_complexity:.2f})")

    def device_levels(self, coarse_max: int = COARSEST_UNKNOWNS,
                      wide_classes: bool = False) -> list[dict]:
        """The hierarchy in the form ``pcfd_set_pressure_solver`` takes.

        The device solves the first level with at most ``coarse_max`` unknowns
        exactly (its dense pseudo-inverse) and drops the levels below it;
        ``coarse_max = COARSEST_UNKNOWNS`` keeps the reference hierarchy whole.
        ``wide_classes`` hands every stored level over with the 32-bit class
        index that levels past 65535 distinct rows need anyway.
        """
        with self._lock:
            self._ensure_depth(coarse_max)
            m = len(self.levels) - 1
            for l, lv in enumerate(self.levels):
                if lv.n <= coarse_max:
                    m = l
                    break
            key = (m, bool(wide_classes))
            if key in self._device:
                return self._device[key]
            out = []
            for l, lv in enumerate(self.levels[:m + 1]):
                last = l == m
                fx, fy = (1, 1) if last else lv.factors
                d = {"nx": lv.shape[0], "ny": lv.shape[1], "fx": fx, "fy": fy,
                     "weight": lv.weight, "active": lv._get("active")}
                if l > 0:
                    g = self._built.copy(l, "di", "dj", "cls", "table", "class_dinv",
                                         "class_active")
                    ncls = g["class_dinv"].size
                    wide = wide_classes or ncls > 65535
                    d.update(di=g["di"], dj=g["dj"],
                             cls=g["cls"] if wide else g["cls"].astype(np.uint16),
                             table=g["table"].reshape(ncls, g["di"].size),
                             class_dinv=g["class_dinv"], class_active=g["class_active"])
                if last:
                    if lv is not self.levels[-1] and lv.active.any():
                        lv.pin = int(np.flatnonzero(lv.active)[0])
                    d["coarse_inv"] = coarsest_pseudo_inverse(lv)
                out.append(d)
            self._device[key] = out
            return out

    def stencil_sizes(self) -> list[tuple[int, int]]:
        """(offsets, distinct rows) per coarse level."""
        return [(int(d["di"].size), int(d["class_dinv"].size)) for d in self.device_levels()[1:]]


# Hierarchies already built in this process, by operator: many simulations of
# one grid and boundary set (a batch of same-size cavities, a run repeated in
# another precision) share one operator, so they share its hierarchy, which is
# read-only once built.  An entry lives as long as some solver holds it.
_SHARED: "weakref.WeakValueDictionary[bytes, MultigridHierarchy]" = weakref.WeakValueDictionary()
_SHARED_LOCK = threading.Lock()


def _digest():
    try:
        import xxhash
        return xxhash.xxh3_128()
    except ImportError:  # pragma: no cover - the image ships xxhash
        return hashlib.blake2b(digest_size=32)


def _operator_key(A: sp.csr_matrix, shape, active, singular, sweeps, periodic) -> bytes:
    h = _digest()
    A = sp.csr_matrix(A)
    for arr in (A.indptr, A.indices, A.data, np.asarray(active, dtype=bool)):
        h.update(np.ascontiguousarray(arr).view(np.uint8))
    h.update(repr((tuple(shape), bool(singular), int(sweeps), tuple(map(bool, periodic)),
                   A.shape)).encode())
    return h.digest()


def shared_hierarchy(A: sp.csr_matrix, shape, active, singular: bool, sweeps: int = 1,
                     periodic: tuple[bool, bool] = (False, False),
                     coarse_max: int = COARSEST_UNKNOWNS) -> "MultigridHierarchy":
    """The hierarchy of this operator, built once per process while in use
    (the same object for every solver of an identical operator; one built
    less deep is deepened when a solver needs a coarser device level)."""
    key = _operator_key(A, shape, active, singular, sweeps, periodic)
    with _SHARED_LOCK:
        hit = _SHARED.get(key)
    if hit is not None:
        return hit
    built = MultigridHierarchy(A, shape, active, singular, presmooth=sweeps,
                               postsmooth=sweeps, periodic=periodic, coarse_max=coarse_max)
    with _SHARED_LOCK:
        return _SHARED.setdefault(key, built)
