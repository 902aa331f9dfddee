"""Batched and served modes: several independent point sets on one device.

``Engine`` (vo_engine_*), ``batch_triangulate`` (vo_batch_triangulate) and ``Server``
(vo_server_*) of include/voronoids.h.  Every set gives the same simplices as
``DelaunayTree(p)`` + ``add_points_to_tree(p, seed)`` on its own points (the reference's
DelaunayTree::new + add_points_to_tree, src/delaunay_tree.rs:389-510 / :544-640 and :336-386);
results are ``SetResult`` records whose ``simplices`` use the reference id space (input point
i = 2(d+1) + i).  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


@dataclass
class SetResult:
    tag: int
    dim: int
    status: int
    n_points: int
    points_duplicate: int
    simplices: np.ndarray          # (n_simplices, dim + 1) int32, reference id space
    t_submit: float
    t_admit: float
    t_done: float
    rounds: int

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit

    def canonical(self) -> tuple[np.ndarray, np.ndarray]:
        """canonical T (all-real simplices as sorted input indices, lexicographic) and E
        (real-real pairs, sorted, unique): the same form as DelaunayTree.canonical()"""
        return canonical_from_simplices(self.simplices, self.dim)


def canonical_from_simplices(simplices: np.ndarray, dim: int) -> tuple[np.ndarray, np.ndarray]:
    M = dim + 1
    s = np.ascontiguousarray(simplices, dtype=np.int32).reshape(-1, M)
    r = _lib.VoSetResult()
    r.dim = dim
    r.n_simplices = s.shape[0]
    r.simplices = s.ctypes.data_as(_i32p)
    T, E = _i32p(), _i32p()
    nT, nE = ctypes.c_int64(), ctypes.c_int64()
    L = _lib.lib()
    _lib.check(L.vo_result_canonical(ctypes.byref(r), ctypes.byref(T), ctypes.byref(nT), ctypes.byref(E),
                                     ctypes.byref(nE)))
    try:
        Ta = np.ctypeslib.as_array(T, shape=(max(nT.value, 1) * M,))[: nT.value * M].copy()
        Ea = np.ctypeslib.as_array(E, shape=(max(nE.value, 1) * 2,))[: nE.value * 2].copy()
    finally:
        L.vo_free(ctypes.cast(T, ctypes.c_void_p))
        L.vo_free(ctypes.cast(E, ctypes.c_void_p))
    return Ta.reshape(-1, M), Ea.reshape(-1, 2)


def plain_simplices(points, seed: int = 0, device: int = 0) -> tuple[np.ndarray, int]:
    """the one-set path on one set: vo_tri_create + vo_tri_insert + vo_tri_export (simplices)
    + vo_tri_destroy; returns the live simplices (reference id space) and the status"""
    p = np.ascontiguousarray(points, dtype=np.float64)
    L = _lib.lib()
    h = ctypes.c_void_p()
    _lib.check(L.vo_tri_create(device, p.shape[1], p.ctypes.data_as(_dp), p.shape[0], ctypes.byref(h)))
    try:
        rc = L.vo_tri_insert(h, p.ctypes.data_as(_dp), p.shape[0], seed)
        if rc not in (_lib.VO_OK, _lib.VO_ERR_DUPLICATE):
            _lib.check(rc)
        return _export_simplices(h, p.shape[1]), rc
    finally:
        L.vo_tri_destroy(h)


def _export_simplices(h, dim: int) -> np.ndarray:
    L = _lib.lib()
    ns = ctypes.c_int64()
    _lib.check(L.vo_tri_counts(h, None, ctypes.byref(ns), None))
    s = np.zeros((ns.value, dim + 1), np.int32)
    _lib.check(L.vo_tri_export(h, s.ctypes.data_as(_i32p), None))
    return s


class PlainPath:
    """The one-set path with one handle per dimension, made into each new set's tree with
    vo_tri_reset (no device allocation per set): insert, export the simplices."""

    def __init__(self, device: int = 0):
        self.device = device
        self.h = {}

    def simplices(self, points, seed: int = 0) -> tuple[np.ndarray, int]:
        p = np.ascontiguousarray(points, dtype=np.float64)
        dim = p.shape[1]
        L = _lib.lib()
        h = self.h.get(dim)
        if h is None:
            h = ctypes.c_void_p()
            _lib.check(L.vo_tri_create(self.device, dim, p.ctypes.data_as(_dp), p.shape[0], ctypes.byref(h)))
            self.h[dim] = h
        else:
            _lib.check(L.vo_tri_reset(h, p.ctypes.data_as(_dp), p.shape[0]))
        rc = L.vo_tri_insert(h, p.ctypes.data_as(_dp), p.shape[0], seed)
        if rc not in (_lib.VO_OK, _lib.VO_ERR_DUPLICATE):
            _lib.check(rc)
        return _export_simplices(h, dim), rc

    def close(self) -> None:
        for h in self.h.values():
            _lib.lib().vo_tri_destroy(h)
        self.h = {}


def simplex_set_hash(simplices) -> int:
    """order-independent 64-bit hash of a set of simplices (vertex tuples sorted, each mixed,
    summed mod 2^64): equal simplex sets give equal hashes whatever their order"""
    s = np.sort(np.asarray(simplices, dtype=np.int64), axis=1).astype(np.uint64)
    h = np.zeros(s.shape[0], np.uint64)
    with np.errstate(over="ignore"):
        for k in range(s.shape[1]):
            h = (h ^ s[:, k]) * np.uint64(0x9E3779B97F4A7C15)
            h ^= h >> np.uint64(29)
        return int(np.sum(h, dtype=np.uint64))


def _take(r: _lib.VoSetResult) -> SetResult:
    """wrap a C result's buffer as a numpy array without copying; vo_free runs when the array
    is collected"""
    M = r.dim + 1
    n = r.n_simplices
    addr = ctypes.cast(r.simplices, ctypes.c_void_p).value if r.simplices else None
    if n > 0 and addr:
        buf = (ctypes.c_int32 * (n * M)).from_address(addr)
        weakref.finalize(buf, _lib.lib().vo_free, ctypes.c_void_p(addr))
        s = np.frombuffer(buf, dtype=np.int32).reshape(n, M)
    else:
        s = np.zeros((0, M), np.int32)
        if addr:
            _lib.lib().vo_free(ctypes.c_void_p(addr))
    r.simplices = _i32p()
    return SetResult(r.tag, r.dim, r.status, r.n_points, r.points_duplicate, s, r.t_submit, r.t_admit, r.t_done,
                     r.rounds)


def clock() -> float:
    """the engines' clock (seconds, steady)"""
    return _lib.lib().vo_clock()


class Engine:
    """vo_engine: sets of one dimension sharing one device arena"""

    def __init__(self, dim: int, max_points: int = 1 << 24, max_simplices: int = 0, device: int = 0, **config):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().vo_engine_create(device, dim, int(max_points), int(max_simplices), ctypes.byref(h)))
        self._h = h
        for k, v in config.items():
            self.config(k, v)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().vo_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self, key: str, value: int) -> None:
        _lib.check(_lib.lib().vo_engine_config(self._h, key.encode(), int(value)))

    def submit(self, points, seed: int = 0, tag: int = 0) -> None:
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, self.dim)
        _lib.check(_lib.lib().vo_engine_submit(self._h, p.ctypes.data_as(_dp), p.shape[0], seed, tag))

    def step(self) -> int:
        left = ctypes.c_int64()
        _lib.check(_lib.lib().vo_engine_step(self._h, ctypes.byref(left)))
        return left.value

    def results(self) -> list[SetResult]:
        out = []
        buf = (_lib.VoSetResult * 256)()
        n = ctypes.c_int64()
        while True:
            _lib.check(_lib.lib().vo_engine_results(self._h, buf, 256, ctypes.byref(n)))
            out.extend(_take(buf[i]) for i in range(n.value))
            if n.value < 256:
                return out

    def run(self) -> list[SetResult]:
        """step until every submitted set is done; results in completion order"""
        out = []
        while self.step() > 0:
            out.extend(self.results())
        out.extend(self.results())
        return out

    def stats(self) -> dict:
        s = _lib.VoEngineStats()
        _lib.check(_lib.lib().vo_engine_stats_get(self._h, ctypes.byref(s)))
        return s.as_dict()

    def batch(self, sets, seed: int = 0) -> list[SetResult]:
        """vo_batch_triangulate: every set to completion, results in set order (seed of set k
        = seed + k)"""
        arrs = [np.ascontiguousarray(s, dtype=np.float64).reshape(-1, self.dim) for s in sets]
        offs = np.zeros(len(arrs) + 1, np.int64)
        offs[1:] = np.cumsum([a.shape[0] for a in arrs])
        pts = np.concatenate(arrs) if arrs else np.zeros((0, self.dim))
        return self.batch_flat(pts, offs, seed)

    def batch_flat(self, pts, offsets, seed: int = 0) -> list[SetResult]:
        """the same over concatenated points: set k = rows offsets[k] .. offsets[k+1]-1"""
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, self.dim)
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        n_sets = offs.shape[0] - 1
        res = (_lib.VoSetResult * max(n_sets, 1))()
        rc = _lib.lib().vo_batch_triangulate(self._h, pts.ctypes.data_as(_dp), offs.ctypes.data_as(_i64p), n_sets,
                                             seed, res)
        out = [_take(res[i]) for i in range(n_sets)]
        _lib.check(rc)
        return out


def batch_triangulate(sets, dim: int | None = None, seed: int = 0, device: int = 0, **config) -> list[SetResult]:
    """many independent sets of one dimension in one engine (sized to the batch)"""
    sets = [np.asarray(s, dtype=np.float64) for s in sets]
    if dim is None:
        dim = sets[0].shape[1]
    biggest = max((s.shape[0] for s in sets), default=0)
    total = sum(s.shape[0] for s in sets)
    max_points = int(min(max(total + (dim + 1) * len(sets), 1 << 16), max(1 << 24, 4 * biggest)))
    eng = Engine(dim, max_points=max_points, device=device, **config)
    try:
        return eng.batch(sets, seed=seed)
    finally:
        eng.close()


class Server:
    """vo_server: continuous batching of 2D and 3D requests (a driver thread in the library)"""

    def __init__(self, max_points_2d: int = 1 << 24, max_points_3d: int = 1 << 24, device: int = 0):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().vo_server_create(device, int(max_points_2d), int(max_points_3d), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().vo_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, points, seed: int = 0, tag: int = 0) -> None:
        p = np.ascontiguousarray(points, dtype=np.float64)
        if p.ndim != 2 or p.shape[1] not in (2, 3):
            raise ValueError("points must be (n, 2) or (n, 3)")
        _lib.check(_lib.lib().vo_server_submit(self._h, p.shape[1], p.ctypes.data_as(_dp), p.shape[0], seed, tag))

    def results(self, timeout_ms: int = 0) -> list[SetResult]:
        buf = (_lib.VoSetResult * 256)()
        n = ctypes.c_int64()
        _lib.check(_lib.lib().vo_server_results(self._h, buf, 256, ctypes.byref(n), int(timeout_ms)))
        return [_take(buf[i]) for i in range(n.value)]

    def stats(self, dim: int) -> dict:
        s = _lib.VoEngineStats()
        _lib.check(_lib.lib().vo_server_stats_get(self._h, dim, ctypes.byref(s)))
        return s.as_dict()


PRED_TYPES = {"orient2d": (1, 2, 3), "incircle": (2, 2, 4), "orient3d": (3, 3, 4), "insphere": (4, 3, 5)}


def predicates(name: str, pts, center=None, device: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """signs and deciding tiers (1 FP32, 2 FP64 filter, 3 FP64 expansions) of predicate
    instances through the construction's cascade (vo_predicates); pts: (n, arity, dim)"""
    t, dim, arity = PRED_TYPES[name]
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, arity, dim)
    n = p.shape[0]
    if center is None:
        flat = p.reshape(-1, dim)
        center = (flat.min(axis=0) + flat.max(axis=0)) / 2 if n else np.zeros(dim)
    c = np.zeros(3)
    c[:dim] = center
    signs = np.zeros(n, np.int32)
    tiers = np.zeros(n, np.int32)
    _lib.check(_lib.lib().vo_predicates(device, t, p.ctypes.data_as(_dp), n, c.ctypes.data_as(_dp),
                                        signs.ctypes.data_as(_i32p), tiers.ctypes.data_as(_i32p)))
    return signs, tiers
