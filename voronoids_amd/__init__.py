"""voronoids_amd: MI355X-native parallel randomized incremental Delaunay construction (2D/3D).

Drop-in for the reference's surface (the ``voronoids`` package re-exports it under the
reference's module name):
  * ``DelaunayTree(points)`` == ``DelaunayTree::new`` (src/delaunay_tree.rs:389-510, :544-640):
    bounding sphere, radius x10, super simplex; no point inserted;
  * ``DelaunayTree.add_points_to_tree(points)`` == ``add_points_to_tree`` (:336-386);
  * ``DelaunayTree.check_delaunay()`` == ``check_delaunay`` (:512-541, :642-671);
  * ``delaunay(points)`` == ``voronoids.delaunay`` (src/lib.rs:104-125) returning a
    ``PyDelauanyTree`` with ``max_simplex_id``, ``vertices`` and ``simplices`` (src/lib.rs:62-102).

All construction runs on the GPU through libvoronoids_hip.so (include/voronoids.h); there is
no CPU fallback.  The views hold what the reference's maps hold: the super vertices 0..d, the
ghost vertices d+1..2d+1, the ghost simplices 1..d+1 (src/delaunay_tree.rs:467-502, :606-632)
with their incidence and adjacency, and every live simplex, super-incident ones included.
Differences from the reference, by design: a live simplex's id is d + 2 + its position in the
export (not its creation id; the super simplex before the first insertion keeps id 0), its
neighbor list is positional (entry k across the face opposite vertex k), and ``center`` /
``radius`` are computed when the view is built, with the reference's own formulas
(src/geometry.rs:2-56, ``vo_circumspheres``), instead of being stored.
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


def _ghosts(dim: int) -> list:
    """ghost simplices 1..d+1 of DelaunayTree::new (src/delaunay_tree.rs:467-502 (3D), :606-632 (2D)):
    ghost vertex first, then the super-hull face it guards"""
    if dim == 3:
        return [[4, 0, 1, 2], [5, 0, 2, 3], [6, 0, 3, 1], [7, 1, 2, 3]]
    return [[3, 0, 1], [4, 0, 2], [5, 1, 2]]


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
    def __init__(self, vertices, device: int = 0, tuned: bool = True):
        """DelaunayTree::new (src/delaunay_tree.rs:389-510, :544-640) on HIP device `device`;
        tuned: apply the autotuner's stored runtime setting for this dimension (tuning.json)"""
        p = _as_points(vertices)
        self.dim = p.shape[1]
        self.M = self.dim + 1
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().vo_tri_create(device, self.dim, p.ctypes.data_as(_dp), p.shape[0], ctypes.byref(h)))
        self._h = h
        self._n_input = 0
        # the sphere DelaunayTree::new stores on simplex 0: (bounding-sphere center, 10 x radius)
        # (an empty or single-point input, which has no super simplex in the reference, uses radius 1)
        c = np.zeros(3)
        r = ctypes.c_double(0.0)
        if p.shape[0]:
            _lib.check(_lib.lib().vo_bounding_sphere(self.dim, p.ctypes.data_as(_dp), p.shape[0],
                                                     c.ctypes.data_as(_dp), ctypes.byref(r)))
        self._root_sphere = (c[:self.dim].tolist(), 10.0 * (r.value if r.value > 0 else 1.0))
        if tuned:
            for key, value in _tuned(self.dim).items():
                self.config(key, value)

    def reset(self, vertices) -> None:
        """the same handle as DelaunayTree(vertices) again, its device buffers kept (vo_tri_reset)"""
        p = _as_points(vertices, self.dim)
        _lib.check(_lib.lib().vo_tri_reset(self._h, p.ctypes.data_as(_dp), p.shape[0]))
        self._n_input = 0
        c = np.zeros(3)
        r = ctypes.c_double(0.0)
        if p.shape[0]:
            _lib.check(_lib.lib().vo_bounding_sphere(self.dim, p.ctypes.data_as(_dp), p.shape[0],
                                                     c.ctypes.data_as(_dp), ctypes.byref(r)))
        self._root_sphere = (c[:self.dim].tolist(), 10.0 * (r.value if r.value > 0 else 1.0))

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

    def locate(self, point) -> tuple[np.ndarray, np.ndarray]:
        """locate + get_new_simplices (src/delaunay_tree.rs:33-123) against the current tree,
        nothing changed (vo_tri_locate): the cavity as sorted export positions, and one row per
        boundary facet: (cavity simplex, simplex across or -1, the facet's vertex ids)"""
        q = np.ascontiguousarray(point, dtype=np.float64)
        if q.shape != (self.dim,):
            raise ValueError(f"point must have {self.dim} coordinates")
        L = _lib.lib()
        nk, nf = ctypes.c_int64(), ctypes.c_int64()
        cap_k, cap_f = 64, 128
        while True:
            k = np.zeros(cap_k, np.int32)
            f = np.zeros((cap_f, self.M + 1), np.int32)
            rc = L.vo_tri_locate(self._h, q.ctypes.data_as(_dp), k.ctypes.data_as(_i32p), cap_k, ctypes.byref(nk),
                                 f.ctypes.data_as(_i32p), cap_f, ctypes.byref(nf))
            if rc == _lib.VO_ERR_OVERFLOW and (nk.value > cap_k or nf.value > cap_f):
                cap_k, cap_f = max(cap_k, nk.value), max(cap_f, nf.value)
                continue
            _lib.check(rc)
            return k[: nk.value], f[: nf.value]

    def check_pairs(self) -> int:
        """the reference's own check_delaunay count (vo_tri_check_pairs): (all-real simplex, vertex)
        pairs with the vertex strictly inside the sphere; O(simplices x vertices)"""
        n = ctypes.c_int64()
        _lib.check(_lib.lib().vo_tri_check_pairs(self._h, ctypes.byref(n)))
        return n.value

    def set_vertex(self, vertex_id: int, point, records: bool = True) -> None:
        """move an inserted vertex without re-triangulating (the reference's pub
        Vertex::coordinates); records=False leaves the simplex records' copies stale"""
        q = np.ascontiguousarray(point, dtype=np.float64)
        if q.shape != (self.dim,):
            raise ValueError(f"point must have {self.dim} coordinates")
        _lib.check(_lib.lib().vo_tri_set_vertex(self._h, int(vertex_id), q.ctypes.data_as(_dp), int(records)))

    def swap_simplex_vertices(self, simplex: int, i: int, j: int) -> None:
        """swap vertex entries i and j of the live simplex at export position `simplex` (the
        reference's pub Simplex::vertices); its neighbor entries stay"""
        _lib.check(_lib.lib().vo_tri_swap_simplex_vertices(self._h, int(simplex), int(i), int(j)))

    def _views(self):
        """ids, vertex rows, neighbor-id rows of live simplices, the ghost rows, coordinates"""
        d, M = self.dim, self.M
        s, nb = self.export()
        X = self.coordinates()
        n = s.shape[0]
        ids = (d + 2) + np.arange(n, dtype=np.int64)
        root = np.all(s < M, axis=1)               # the super simplex, live until the first insertion
        ids[root] = 0
        ghosts = _ghosts(d)
        face_ghost = {frozenset(g[1:]): k + 1 for k, g in enumerate(ghosts)}
        nbr_ids = np.where(nb >= 0, ids[np.maximum(nb, 0)], -1)
        ghost_nbr = [[] for _ in ghosts]
        for i, f in zip(*np.nonzero(nb < 0)):
            g = face_ghost[frozenset(int(x) for k, x in enumerate(s[i]) if k != f)]
            nbr_ids[i, f] = g
            ghost_nbr[g - 1].append(int(ids[i]))
        return ids, s, nbr_ids, ghosts, ghost_nbr, X

    @property
    def simplices(self) -> dict:
        ids, s, nbr_ids, ghosts, ghost_nbr, X = self._views()
        n, d = s.shape[0], self.dim
        center = np.zeros((n, d))
        radius = np.zeros(n)
        if n:
            _lib.check(_lib.lib().vo_circumspheres(d, X.ctypes.data_as(_dp), X.shape[0],
                                                   np.ascontiguousarray(s).ctypes.data_as(_i32p), n,
                                                   center.ctypes.data_as(_dp), radius.ctypes.data_as(_dp)))
        out = {}
        for k, g in enumerate(ghosts):              # center 0, radius 0: never in conflict
            out[k + 1] = PySimplex(list(g), [0.0] * d, 0.0, ghost_nbr[k])
        for i in range(n):
            c, r = center[i].tolist(), float(radius[i])
            if ids[i] == 0:                          # simplex 0 keeps the sphere new() gave it
                c, r = list(self._root_sphere[0]), float(self._root_sphere[1])
            out[int(ids[i])] = PySimplex(s[i].tolist(), c, r, nbr_ids[i].tolist())
        return out

    @property
    def vertices(self) -> dict:
        ids, s, _, ghosts, _, X = self._views()
        M = self.M
        flat = np.concatenate([s.ravel().astype(np.int64), np.asarray(ghosts, np.int64).ravel()])
        owner = np.concatenate([np.repeat(ids, M), np.repeat(np.arange(1, M + 1, dtype=np.int64), M)])
        order = np.lexsort((owner, flat))
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
