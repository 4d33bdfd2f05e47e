"""Aggregation-multigrid hierarchy: host setup, device V-cycle.

The hierarchy is built once per operator, as in ``pycfd/core/multigrid.py``:
geometric 2x2 / 2x1 / 1x2 aggregates that keep each level a rectangular grid
with holes (:167-245), the 0/1 tentative prolongator smoothed by one damped
Jacobi pass (:248-273), Galerkin coarse operators with the identity on
dropped aggregates (:276-287, 380-383), damping weights from a seeded
12-step power iteration (:129-157, 354-362), and the coarsest level solved
exactly as a pinned solve followed by mean removal (:514-543).

Because every level is a structured grid, each coarse operator is handed to
the GPU as a row dictionary: a per-cell class index into the level's distinct
operator rows (interior rows repeat by translation invariance), and the
prolongator is never stored: the device applies ``P = (I - w D^-1 A) P_tent``
and ``P^T`` on the fly from the level's own operator.  The device's coarsest
level is the first one with at most ``coarse_max`` unknowns, solved by its
dense FP64 pseudo-inverse; with ``coarse_max`` at the reference's 96 this is
the reference's own coarsest level.
"""

from __future__ import annotations

import hashlib
import logging
import os
import threading
import weakref
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

log = logging.getLogger(__name__)

COARSEST_UNKNOWNS = 96
MAX_LEVELS = 20
MIN_COARSENING_RATIO = 1.3
JACOBI_WEIGHT_FACTOR = 4.0 / 3.0
POWER_ITERATIONS = 12
STALL_FACTOR = 0.995
STALL_PATIENCE = 3
ANISOTROPY_THRESHOLD = 2.0


# Sparse products are split into row blocks run on threads (SciPy's sparse
# kernels release the interpreter lock).  Each output row is computed by one
# block exactly as the whole product computes it, so the result does not
# depend on the split.
_SETUP_THREADS = max(1, min(8, os.cpu_count() or 1))
_BLOCKED_MIN_NNZ = 200_000
_POOL: ThreadPoolExecutor | None = None


def _pool() -> ThreadPoolExecutor:
    global _POOL
    if _POOL is None:
        _POOL = ThreadPoolExecutor(_SETUP_THREADS, thread_name_prefix="pcfd-mg-setup")
    return _POOL


def _row_cuts(X: sp.csr_matrix) -> np.ndarray:
    """Row block boundaries with about equal nonzeros per block."""
    n = X.shape[0]
    targets = np.linspace(0, X.nnz, _SETUP_THREADS + 1)
    cuts = np.searchsorted(X.indptr, targets).clip(0, n)
    cuts[0], cuts[-1] = 0, n
    return np.unique(cuts)


def _rows_view(X: sp.csr_matrix, a: int, b: int) -> sp.csr_matrix:
    """Rows [a, b) of X sharing X's data and column arrays (no copy)."""
    s, e = X.indptr[a], X.indptr[b]
    return sp.csr_matrix((X.data[s:e], X.indices[s:e], X.indptr[a:b + 1] - s),
                         shape=(b - a, X.shape[1]), copy=False)


def _matmul(X: sp.csr_matrix, Y: sp.csr_matrix) -> sp.csr_matrix:
    """X @ Y (both CSR), by row blocks of X on threads."""
    X, Y = sp.csr_matrix(X), sp.csr_matrix(Y)
    if _SETUP_THREADS == 1 or X.nnz < _BLOCKED_MIN_NNZ:
        return sp.csr_matrix(X @ Y)
    cuts = _row_cuts(X)
    parts = list(_pool().map(lambda ab: sp.csr_matrix(_rows_view(X, ab[0], ab[1]) @ Y),
                             zip(cuts[:-1], cuts[1:])))
    offs = np.cumsum([0] + [q.nnz for q in parts])
    indptr = np.concatenate([parts[0].indptr[:1]] +
                            [q.indptr[1:] + o for q, o in zip(parts, offs[:-1])])
    return sp.csr_matrix((np.concatenate([q.data for q in parts]),
                          np.concatenate([q.indices for q in parts]), indptr),
                         shape=(X.shape[0], Y.shape[1]), copy=False)


def spectral_radius(A: sp.csr_matrix, dinv: np.ndarray, seed: int = 0) -> float:
    """Power-iteration estimate of rho(D^-1 A), nudged up 5% (multigrid.py:129-157)."""
    n = A.shape[0]
    if n == 0:
        return 1.0
    x = np.random.default_rng(seed).standard_normal(n)
    nrm = float(np.linalg.norm(x))
    if nrm == 0.0:
        return 1.0
    x /= nrm
    rho = 1.0
    for _ in range(POWER_ITERATIONS):
        y = dinv * (A @ x)
        nrm = float(np.linalg.norm(y))
        if nrm == 0.0:
            return rho
        rho = nrm
        x = y / nrm
    return rho * 1.05


def coarsening_factors(A: sp.csr_matrix, shape: tuple[int, int]) -> tuple[int, int]:
    """Full or semi-coarsening from the operator's axis strengths (multigrid.py:167-204)."""
    nx, ny = shape
    if nx < 2 and ny < 2:
        return 1, 1

    def strength(offset: int) -> float:
        if offset >= A.shape[0]:
            return 0.0
        band = np.abs(A.diagonal(offset))
        return float(band.mean()) if band.size else 0.0

    sx, sy = strength(ny), strength(1)
    if nx >= 2 and sx > ANISOTROPY_THRESHOLD * sy:
        return 2, 1
    if ny >= 2 and sy > ANISOTROPY_THRESHOLD * sx:
        return 1, 2
    return (2 if nx >= 2 else 1), (2 if ny >= 2 else 1)


def coarse_shape(shape, factors) -> tuple[int, int]:
    return tuple(max(-(-n // f), 1) for n, f in zip(shape, factors))  # type: ignore[return-value]


def aggregate_blocks(shape, active: np.ndarray, factors):
    """Block labels (-1 on inactive cells) and the coarse active mask (multigrid.py:212-245)."""
    nx, ny = shape
    fx, fy = factors
    ncx, ncy = coarse_shape(shape, factors)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    block = ((i // fx) * ncy + (j // fy)).ravel()
    labels = np.where(np.asarray(active).ravel(), block, -1)
    coarse_active = np.zeros(ncx * ncy, dtype=bool)
    coarse_active[labels[labels >= 0]] = True
    return labels, coarse_active


def tentative_prolongator(labels: np.ndarray, n_coarse: int) -> sp.csr_matrix:
    rows = np.flatnonzero(labels >= 0)
    return sp.coo_matrix((np.ones(rows.size), (rows, labels[rows])),
                         shape=(labels.size, n_coarse)).tocsr()


def smooth_prolongator(A, dinv, weight, P) -> sp.csr_matrix:
    return (P - sp.diags(weight * dinv) @ _matmul(A, P)).tocsr()


def _identity_on_inactive(A: sp.csr_matrix, active: np.ndarray) -> sp.csr_matrix:
    if active.all():
        return A
    return (A + sp.diags((~active).astype(float))).tocsr()


@dataclass
class Level:
    A: sp.csr_matrix
    dinv: np.ndarray
    weight: float
    active: np.ndarray
    shape: tuple[int, int]
    singular: bool
    factors: tuple[int, int] = (1, 1)   # towards the next coarser level
    pin: int = 0

    @property
    def n(self) -> int:
        return self.A.shape[0]


def _distinct_rows(rows: np.ndarray):
    """``np.unique(rows, axis=0, return_inverse=True)`` for float rows, by
    a 64-bit hash of each row's bit pattern: one sort of n hashes instead of a
    lexicographic sort of n rows.  Rows that share a hash are verified equal
    bit for bit, so a collision cannot merge distinct rows (it falls back to
    the exact sort).  Returns the distinct rows in ascending lexicographic
    order, as np.unique does, and each row's index into them."""
    n, w = rows.shape
    bits = np.ascontiguousarray(rows).view(np.uint64)
    # +0.0 and -0.0 compare equal, as np.unique treats them
    bits = np.where(bits == np.uint64(1 << 63), np.uint64(0), bits)
    mult = np.random.default_rng(12345).integers(1, 2 ** 63, size=w, dtype=np.uint64) | np.uint64(1)
    with np.errstate(over="ignore"):
        h = (bits * mult[None, :]).sum(axis=1, dtype=np.uint64)
    uh, first, inv = np.unique(h, return_index=True, return_inverse=True)
    reps = bits[first]
    if not np.array_equal(bits, reps[inv]):  # a hash shared by two different rows
        return np.unique(rows, axis=0, return_inverse=True)
    table = rows[first]
    order = np.lexsort(table.T[::-1])
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return table[order], rank[inv.ravel()]


def to_class_stencil(A: sp.csr_matrix, shape: tuple[int, int], periodic: tuple[bool, bool],
                     active: np.ndarray, wide: bool = False):
    """CSR operator on a rectangular grid -> stencil dictionary.

    Returns offsets ``(di, dj)``, the per-cell class index, the class table
    (one row of coefficients per distinct operator row) and, per class, the
    diagonal inverse and the active flag.  The class index is ``uint16``, or
    ``uint32`` past 65535 classes (a stretched mesh, whose rows all differ)
    or when ``wide`` asks for it.
    """
    nx, ny = shape
    coo = A.tocoo()
    r, c, v = coo.row, coo.col, coo.data
    keep = v != 0.0
    r, c, v = r[keep], c[keep], v[keep]
    ri, rj = np.divmod(r, ny)
    ci, cj = np.divmod(c, ny)
    di, dj = ci - ri, cj - rj
    if periodic[0] and nx > 1:
        di = (di + nx // 2) % nx - nx // 2
    if periodic[1] and ny > 1:
        dj = (dj + ny // 2) % ny - ny // 2
    if np.abs(di).max(initial=0) > 127 or np.abs(dj).max(initial=0) > 127:
        raise ValueError("coarse stencil offset exceeds the int8 range")
    # distinct offsets in (di, dj) order and each entry's index among them, by
    # a presence table over the int8 offset range (no sort of the entries)
    key = (di.astype(np.int64) + 128) * 256 + (dj.astype(np.int64) + 128)
    present = np.zeros(256 * 256, dtype=bool)
    present[key] = True
    uniq = np.flatnonzero(present)
    slot = np.cumsum(present) - 1
    inv = slot[key]
    udi = uniq // 256 - 128
    udj = uniq % 256 - 128
    S = uniq.size
    rows = np.zeros((nx * ny, S + 1))
    rows[r, inv] = v
    rows[:, S] = np.asarray(active, dtype=float).ravel()
    table, cls = _distinct_rows(rows)
    if table.shape[0] >= 2 ** 32:
        raise ValueError(f"{table.shape[0]} distinct operator rows exceed the class index range")
    wide = wide or table.shape[0] > 65535
    centre = int(np.flatnonzero((udi == 0) & (udj == 0))[0])
    diag = table[:, centre]
    cdinv = 1.0 / np.where(diag == 0.0, 1.0, diag)
    cactive = table[:, S] != 0.0
    return (udi.astype(np.int8), udj.astype(np.int8),
            cls.ravel().astype(np.uint32 if wide else np.uint16),
            np.ascontiguousarray(table[:, :S]), cdinv, cactive.astype(np.uint8))


def coarsest_pseudo_inverse(level: Level) -> np.ndarray:
    """Dense operator of the coarsest solve, projections included (multigrid.py:514-543).

    Singular case: ``Pi inv(M) Z Pi`` with M the operator pinned at ``pin``,
    Z zeroing the pinned right-hand side entry and Pi the mean removal on the
    active set (zero on inactive cells); the two projections are applied as
    column and row mean subtractions.  Otherwise inactive rows are zeroed.
    """
    n = level.n
    A = level.A.toarray()
    act = level.active
    if level.singular and n:
        M = A.copy()
        M[level.pin, :] = 0.0
        M[level.pin, level.pin] = 1.0
        Y = np.linalg.inv(M)
        Y[:, level.pin] = 0.0
        ia = np.flatnonzero(act)
        out = np.zeros((n, n))
        if ia.size:
            Yaa = Y[np.ix_(ia, ia)]
            Yaa -= Yaa.mean(axis=0, keepdims=True)
            Yaa -= Yaa.mean(axis=1, keepdims=True)
            out[np.ix_(ia, ia)] = Yaa
        return out
    zero_inactive = np.diag(act.astype(float))
    return zero_inactive @ np.linalg.inv(A)


class MultigridHierarchy:
    """A V-cycle hierarchy for one operator (multigrid.py:324-474)."""

    def __init__(self, A: sp.csr_matrix, shape: tuple[int, int], active: np.ndarray,
                 singular: bool, presmooth: int = 1, postsmooth: int = 1,
                 periodic: tuple[bool, bool] = (False, False)) -> None:
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
        self.levels: list[Level] = []
        self._build(sp.csr_matrix(A), tuple(shape), np.asarray(active, dtype=bool).ravel(),
                    bool(singular))
        self._device: dict[tuple[int, bool], list[dict]] = {}

    def _make_level(self, A, shape, active, singular) -> Level:
        diag = A.diagonal()
        dinv = 1.0 / np.where(diag == 0.0, 1.0, diag)
        rho = spectral_radius(A, dinv)
        weight = JACOBI_WEIGHT_FACTOR / rho if rho > 0.0 else 1.0
        return Level(A=A, dinv=dinv, weight=weight, active=active, shape=shape,
                     singular=singular)

    def _build(self, A, shape, active, singular) -> None:
        level = self._make_level(A, shape, active, singular)
        self.levels.append(level)
        while len(self.levels) < MAX_LEVELS and level.n > COARSEST_UNKNOWNS:
            fx, fy = coarsening_factors(level.A, level.shape)
            labels, coarse_active = aggregate_blocks(level.shape, level.active, (fx, fy))
            next_shape = coarse_shape(level.shape, (fx, fy))
            n_coarse = coarse_active.size
            if n_coarse == 0 or level.n < MIN_COARSENING_RATIO * n_coarse:
                break
            P = tentative_prolongator(labels, n_coarse)
            P = smooth_prolongator(level.A, level.dinv, level.weight, P)
            # P^T (A P): both products by row blocks (P^T held as CSR)
            coarse_A = _identity_on_inactive(
                _matmul(sp.csr_matrix(P.T), _matmul(level.A, P)), coarse_active)
            coarse_A.sum_duplicates()
            level.factors = (fx, fy)
            level = self._make_level(coarse_A, next_shape, coarse_active, singular)
            self.levels.append(level)
        last = self.levels[-1]
        last.pin = int(np.flatnonzero(last.active)[0]) if last.active.any() else 0
        log.debug("multigrid hierarchy: %s", self)

    @property
    def n_levels(self) -> int:
        return len(self.levels)

    @property
    def unknowns(self) -> list[int]:
        return [lv.n for lv in self.levels]

    @property
    def operator_complexity(self) -> float:
        fine = self.levels[0].A.nnz
        return sum(lv.A.nnz for lv in self.levels) / fine if fine else 1.0

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
                 "weight": lv.weight, "active": lv.active.astype(np.uint8)}
            if l > 0:
                di, dj, cls, table, cdinv, cact = to_class_stencil(
                    lv.A, lv.shape, self.periodic, lv.active, wide=wide_classes)
                d.update(di=di, dj=dj, cls=cls, table=table, class_dinv=cdinv,
                         class_active=cact)
            if last:
                pin_level = lv
                if lv is not self.levels[-1]:
                    pin_level = Level(A=lv.A, dinv=lv.dinv, weight=lv.weight, active=lv.active,
                                      shape=lv.shape, singular=lv.singular,
                                      pin=int(np.flatnonzero(lv.active)[0])
                                      if lv.active.any() else 0)
                d["coarse_inv"] = coarsest_pseudo_inverse(pin_level)
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


def _operator_key(A: sp.csr_matrix, shape, active, singular, sweeps, periodic) -> bytes:
    h = hashlib.blake2b(digest_size=32)
    A = sp.csr_matrix(A)
    for arr in (A.indptr, A.indices, A.data, np.asarray(active, dtype=bool)):
        h.update(np.ascontiguousarray(arr).view(np.uint8))
    h.update(repr((tuple(shape), bool(singular), int(sweeps), tuple(map(bool, periodic)),
                   A.shape)).encode())
    return h.digest()


def shared_hierarchy(A: sp.csr_matrix, shape, active, singular: bool, sweeps: int = 1,
                     periodic: tuple[bool, bool] = (False, False)) -> "MultigridHierarchy":
    """The hierarchy of this operator, built once per process while in use
    (the same object for every solver of an identical operator)."""
    key = _operator_key(A, shape, active, singular, sweeps, periodic)
    with _SHARED_LOCK:
        hit = _SHARED.get(key)
    if hit is not None:
        return hit
    built = MultigridHierarchy(A, shape, active, singular, presmooth=sweeps,
                               postsmooth=sweeps, periodic=periodic)
    with _SHARED_LOCK:
        return _SHARED.setdefault(key, built)
