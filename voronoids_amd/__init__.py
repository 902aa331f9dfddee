"""voronoids_amd: MI355X-native parallel randomized incremental Delaunay construction (2D/3D).

Drop-in for the reference's surface:
  * ``DelaunayTree(points)`` == ``DelaunayTree::new`` (src/delaunay_tree.rs:389-510, :544-640):
    bounding sphere, radius x10, super simplex; no point inserted;
  * ``DelaunayTree.add_points_to_tree(points)`` == ``add_points_to_tree`` (:336-386);
  * ``DelaunayTree.check_delaunay()`` == ``check_delaunay`` (:512-541, :642-671);
  * ``delaunay(points)`` == ``voronoids.delaunay`` (src/lib.rs:104-125) returning a
    ``PyDelauanyTree`` with ``max_simplex_id``, ``vertices`` and ``simplices`` (src/lib.rs:62-102).

All construction runs on the GPU through libvoronoids_hip.so (include/voronoids.h); there is
no CPU fallback.  Differences from the reference, by design: simplex ids are export
positions, not creation-history ids; hull faces have no ghost simplices (no neighbor);
``center``/``radius`` are computed on export.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import VoError, VoStats
from .batch import Engine, Server, SetResult, batch_triangulate

__all__ = ["DelaunayTree", "PyDelauanyTree", "PyVertex", "PySimplex", "delaunay", "VoError", "Engine", "Server",
           "SetResult", "batch_triangulate"]

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def _as_points(points, dim: int | None = None) -> np.ndarray:
    p = np.ascontiguousarray(points, dtype=np.float64)
    if p.ndim != 2 or p.shape[1] not in (2, 3) or (dim is not None and p.shape[1] != dim):
        raise ValueError("points must be an (n, 2) or (n, 3) array" + (f" of dimension {dim}" if dim else ""))
    return p


@dataclass
class PyVertex:
    """src/lib.rs:12-29"""
    point: list
    simplex: list = field(default_factory=list)


@dataclass
class PySimplex:
    """src/lib.rs:31-60"""
    vertices: list
    center: list
    radius: float
    neighbors: list


def _circumspheres(X: np.ndarray, S: np.ndarray):
    """circumcenters and radii of simplices S (vertex rows) with coordinates X"""
    P = X[S]                                    # (n, M, d)
    A = P[:, 1:, :] - P[:, :1, :]
    b = 0.5 * np.sum(A * A, axis=2)
    try:
        c = np.linalg.solve(A, b[..., None])[..., 0]
    except np.linalg.LinAlgError:
        c = np.full(b.shape, np.nan)
    center = P[:, 0, :] + c
    radius = np.sqrt(np.sum(c * c, axis=1))
    return center, radius


def _tuned(dim: int) -> dict:
    """the autotuner's best round-kernel setting for this dimension (voronoids_amd/tuning.json)"""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get(str(dim), {}).get("setting", {})


class DelaunayTree:
    def __init__(self, vertices, device: int = 0):
        p = _as_points(vertices)
        self.dim = p.shape[1]
        self.M = self.dim + 1
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().vo_tri_create(device, self.dim, p.ctypes.data_as(_dp), p.shape[0], ctypes.byref(h)))
        self._h = h
        self._n_input = 0
        for key, value in _tuned(self.dim).items():
            self.config(key, value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().vo_tri_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self, key: str, value: int) -> None:
        """tuning / test knobs, see vo_tri_config in include/voronoids.h"""
        _lib.check(_lib.lib().vo_tri_config(self._h, key.encode(), int(value)))

    # -- construction -----------------------------------------------------------------
    def add_points_to_tree(self, vertices, seed: int = 0) -> None:
        self.stage(vertices, seed)
        self.run()

    def stage(self, vertices, seed: int = 0) -> None:
        """a failed stage consumes no vertex ids; a successful one consumes one per point, also
        for points the run later rejects as duplicates"""
        p = _as_points(vertices, self.dim)
        _lib.check(_lib.lib().vo_tri_stage(self._h, p.ctypes.data_as(_dp), p.shape[0], seed))
        self._n_input += p.shape[0]

    def run(self) -> None:
        _lib.check(_lib.lib().vo_tri_run(self._h))

    # -- inspection ---------------------------------------------------------------------
    def counts(self) -> tuple[int, int, int]:
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib().vo_tri_counts(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    @property
    def max_simplex_id(self) -> int:
        return self.counts()[2]

    def export(self) -> tuple[np.ndarray, np.ndarray]:
        n_s = self.counts()[1]
        s = np.zeros((n_s, self.M), np.int32)
        nb = np.zeros((n_s, self.M), np.int32)
        _lib.check(_lib.lib().vo_tri_export(self._h, s.ctypes.data_as(_i32p), nb.ctypes.data_as(_i32p)))
        return s, nb

    def coordinates(self) -> np.ndarray:
        X = np.zeros((2 * self.M + self._n_input, self.dim))
        _lib.check(_lib.lib().vo_tri_vertices(self._h, X.ctypes.data_as(_dp)))
        return X

    def canonical(self) -> tuple[np.ndarray, np.ndarray]:
        T, E = _i32p(), _i32p()
        nT, nE = ctypes.c_int64(), ctypes.c_int64()
        L = _lib.lib()
        _lib.check(L.vo_tri_canonical(self._h, ctypes.byref(T), ctypes.byref(nT), ctypes.byref(E), ctypes.byref(nE)))
        try:
            Ta = np.ctypeslib.as_array(T, shape=(max(nT.value, 1) * self.M,))[: nT.value * self.M].copy()
            Ea = np.ctypeslib.as_array(E, shape=(max(nE.value, 1) * 2,))[: nE.value * 2].copy()
        finally:
            L.vo_free(ctypes.cast(T, ctypes.c_void_p))
            L.vo_free(ctypes.cast(E, ctypes.c_void_p))
        return Ta.reshape(-1, self.M), Ea.reshape(-1, 2)

    def violations(self) -> int:
        n = ctypes.c_int64()
        _lib.check(_lib.lib().vo_tri_check(self._h, ctypes.byref(n)))
        return n.value

    def check_delaunay(self) -> bool:
        return self.violations() == 0

    def stats(self) -> dict:
        s = VoStats()
        _lib.check(_lib.lib().vo_tri_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    @property
    def simplices(self) -> dict:
        s, nb = self.export()
        X = self.coordinates()
        center, radius = _circumspheres(X, s.astype(np.int64))
        return {i: PySimplex(s[i].tolist(), center[i].tolist(), float(radius[i]), [int(x) for x in nb[i] if x >= 0])
                for i in range(s.shape[0])}

    @property
    def vertices(self) -> dict:
        s, _ = self.export()
        X = self.coordinates()
        flat = s.ravel().astype(np.int64)
        owner = np.repeat(np.arange(s.shape[0]), self.M)
        order = np.argsort(flat, kind="stable")
        flat, owner = flat[order], owner[order]
        bounds = np.searchsorted(flat, np.arange(X.shape[0] + 1))
        return {v: PyVertex(X[v].tolist(), owner[bounds[v]:bounds[v + 1]].tolist()) for v in range(X.shape[0])}


class PyDelauanyTree(DelaunayTree):
    """the result type of voronoids.delaunay (src/lib.rs:62-102; the typo is the reference's)"""


def delaunay(points, seed: int = 0, device: int = 0) -> PyDelauanyTree:
    """voronoids.delaunay (src/lib.rs:104-125): DelaunayTree::new(points), then every point
    inserted.  Duplicate points raise VoError (the reference panics in locate)."""
    p = _as_points(points)
    t = PyDelauanyTree(p, device=device)
    t.add_points_to_tree(p, seed=seed)
    return t
