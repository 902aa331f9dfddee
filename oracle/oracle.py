"""ctypes wrapper over oracle/liboracle.so, plus the canonical form of a triangulation.

TEST INFRASTRUCTURE ONLY: the oracle is the checker for parity tests and the CPU baseline
of bench.py.  The product (voronoids_amd) never imports this module.

Canonical form (identical on both sides of every parity test):
  T  every live simplex whose vertices are all input points, as input indices, each row
     sorted ascending, rows sorted lexicographically (int32, shape (nT, d+1));
  E  every unordered pair of input points that co-occurs in any live non-ghost simplex,
     super-vertex-incident ones included (int32, shape (nE, 2)), sorted, unique;
  sha256 over T's then E's little-endian int32 bytes.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

MODE_EXACT = 0
MODE_REF = 1

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        for name in ("or_orient2d", "or_orient2d_bigint"):
            getattr(L, name).argtypes = [_dp] * 3
        for name in ("or_orient3d", "or_orient3d_bigint", "or_incircle", "or_incircle_bigint"):
            getattr(L, name).argtypes = [_dp] * 4
        for name in ("or_insphere", "or_insphere_bigint"):
            getattr(L, name).argtypes = [_dp] * 5
        L.or_circumsphere.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.or_bounding_sphere.argtypes = [ctypes.c_int, _dp, ctypes.c_int64, _dp, _dp]
        L.or_super_vertices.argtypes = [ctypes.c_int, _dp, ctypes.c_int64, _dp]
        L.or_trig_constants.argtypes = [_dp]
        L.or_tree_new.argtypes = [ctypes.c_int, _dp, ctypes.c_int64, ctypes.c_int]
        L.or_tree_new.restype = ctypes.c_void_p
        L.or_delaunay.argtypes = [ctypes.c_int, _dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.or_delaunay.restype = ctypes.c_void_p
        L.or_tree_free.argtypes = [ctypes.c_void_p]
        L.or_tree_insert_seq.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64, ctypes.c_int64]
        L.or_tree_add_points_parallel.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64, ctypes.c_int64,
                                                  ctypes.c_int, _i64p, _i64p]
        for name in ("or_tree_max_simplex_id", "or_tree_n_vertices", "or_tree_n_simplices",
                     "or_tree_check_delaunay"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = ctypes.c_int64
        L.or_tree_export.argtypes = [ctypes.c_void_p, _i64p, _i32p, _i32p, _dp, _dp]
        L.or_tree_vertices.argtypes = [ctypes.c_void_p, _i64p, _dp]
        L.or_tree_vertex_simplices.argtypes = [ctypes.c_void_p, ctypes.c_int64, _i64p, ctypes.c_int64]
        L.or_tree_vertex_simplices.restype = ctypes.c_int64
        L.or_find_placement.argtypes = [ctypes.c_int64, _i64p, _i64p, _i64p]
        L.or_tree_set_vertex.argtypes = [ctypes.c_void_p, ctypes.c_int64, _dp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _pt(p) -> np.ndarray:
    return np.ascontiguousarray(p, dtype=np.float64)


def orient2d(a, b, c) -> int:
    return lib().or_orient2d(*[_ptr(_pt(x), _dp) for x in (a, b, c)])


def orient3d(a, b, c, d) -> int:
    return lib().or_orient3d(*[_ptr(_pt(x), _dp) for x in (a, b, c, d)])


def incircle(a, b, c, d) -> int:
    return lib().or_incircle(*[_ptr(_pt(x), _dp) for x in (a, b, c, d)])


def insphere(a, b, c, d, e) -> int:
    return lib().or_insphere(*[_ptr(_pt(x), _dp) for x in (a, b, c, d, e)])


def bigint_sign(name: str, *pts) -> int:
    arrs = [_pt(x) for x in pts]
    return getattr(lib(), f"or_{name}_bigint")(*[_ptr(x, _dp) for x in arrs])


def circumsphere(verts) -> tuple[np.ndarray, float]:
    v = _pt(verts)
    dim = v.shape[1]
    c = np.zeros(dim)
    r = np.zeros(1)
    lib().or_circumsphere(dim, _ptr(v, _dp), _ptr(c, _dp), _ptr(r, _dp))
    return c, float(r[0])


def bounding_sphere(pts) -> tuple[np.ndarray, float]:
    p = _pt(pts)
    dim = p.shape[1]
    c = np.zeros(dim)
    r = np.zeros(1)
    lib().or_bounding_sphere(dim, _ptr(p, _dp), p.shape[0], _ptr(c, _dp), _ptr(r, _dp))
    return c, float(r[0])


def super_vertices(pts) -> np.ndarray:
    """the 2(d+1) super + ghost vertex rows DelaunayTree::new builds"""
    p = _pt(pts)
    dim = p.shape[1]
    out = np.zeros((2 * (dim + 1), dim))
    lib().or_super_vertices(dim, _ptr(p, _dp), p.shape[0], _ptr(out, _dp))
    return out


def trig_constants() -> np.ndarray:
    out = np.zeros(4)
    lib().or_trig_constants(_ptr(out, _dp))
    return out


class Tree:
    """The reference's DelaunayTree restated (src/delaunay_tree.rs:24-30)."""

    def __init__(self, pts, mode: int = MODE_EXACT, _handle=None, dim=None):
        if _handle is not None:
            self.h = _handle
            self.dim = dim
        else:
            p = _pt(pts)
            self.dim = p.shape[1]
            self.h = lib().or_tree_new(self.dim, _ptr(p, _dp), p.shape[0], mode)
        self.M = self.dim + 1

    @classmethod
    def delaunay(cls, pts, mode: int = MODE_EXACT, nthreads: int = 0) -> "Tree":
        """voronoids.delaunay restated (src/lib.rs:104-125), for 2D and 3D"""
        p = _pt(pts)
        h = lib().or_delaunay(p.shape[1], _ptr(p, _dp), p.shape[0], mode, nthreads)
        if not h:
            raise RuntimeError("oracle: locate found no conflicting simplex (duplicate point)")
        return cls(None, _handle=h, dim=p.shape[1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_tree_free(self.h)
            self.h = None

    def insert_seq(self, pts, input_base: int = 0) -> None:
        p = _pt(pts)
        rc = lib().or_tree_insert_seq(self.h, _ptr(p, _dp), p.shape[0], input_base)
        if rc != 0:
            raise RuntimeError("oracle: locate found no conflicting simplex (duplicate point)")

    def add_points_parallel(self, pts, input_base: int = 0, nthreads: int = 0) -> dict:
        p = _pt(pts)
        nl = np.zeros(1, np.int64)
        no = np.zeros(1, np.int64)
        rc = lib().or_tree_add_points_parallel(self.h, _ptr(p, _dp), p.shape[0], input_base, nthreads,
                                               _ptr(nl, _i64p), _ptr(no, _i64p))
        if rc != 0:
            raise RuntimeError("oracle: locate found no conflicting simplex (duplicate point)")
        return {"levels": int(nl[0]), "overlaps": int(no[0])}

    @property
    def max_simplex_id(self) -> int:
        return int(lib().or_tree_max_simplex_id(self.h))

    @property
    def n_vertices(self) -> int:
        return int(lib().or_tree_n_vertices(self.h))

    @property
    def n_simplices(self) -> int:
        return int(lib().or_tree_n_simplices(self.h))

    def export(self) -> dict:
        ns = self.n_simplices
        M, d = self.M, self.dim
        ids = np.zeros(ns, np.int64)
        verts = np.zeros((ns, M), np.int32)
        nbrs = np.zeros((ns, M), np.int32)
        cen = np.zeros((ns, d))
        rad = np.zeros(ns)
        lib().or_tree_export(self.h, _ptr(ids, _i64p), _ptr(verts, _i32p), _ptr(nbrs, _i32p),
                             _ptr(cen, _dp), _ptr(rad, _dp))
        nv = self.n_vertices
        vin = np.zeros(nv, np.int64)
        X = np.zeros((nv, d))
        lib().or_tree_vertices(self.h, _ptr(vin, _i64p), _ptr(X, _dp))
        return {"ids": ids, "verts": verts, "nbrs": nbrs, "centers": cen, "radii": rad,
                "vertex_input": vin, "coords": X}

    def vertex_simplices(self, v: int) -> list[int]:
        buf = np.zeros(256, np.int64)
        n = lib().or_tree_vertex_simplices(self.h, v, _ptr(buf, _i64p), 256)
        if n > 256:
            buf = np.zeros(n, np.int64)
            lib().or_tree_vertex_simplices(self.h, v, _ptr(buf, _i64p), n)
        return buf[:n].tolist()

    def check_delaunay(self) -> int:
        return int(lib().or_tree_check_delaunay(self.h))

    def set_vertex(self, v: int, p) -> None:
        """vertex v moves to p, nothing re-triangulated (the reference's pub Vertex::coordinates)"""
        q = np.ascontiguousarray(p, dtype=np.float64)
        if lib().or_tree_set_vertex(self.h, int(v), _ptr(q, _dp)) != 0:
            raise IndexError("no vertex %d" % v)

    def canonical(self) -> "Canonical":
        ex = self.export()
        return canonical_from_simplices(ex["verts"], ex["vertex_input"], self.dim)


def find_placement(footprints: list[list[int]]) -> np.ndarray:
    n = len(footprints)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(f) for f in footprints])
    fp = np.ascontiguousarray(np.concatenate([np.asarray(f, np.int64) for f in footprints])
                              if n else np.zeros(0, np.int64))
    out = np.zeros(n, np.int64)
    lib().or_find_placement(n, _ptr(off, _i64p), _ptr(fp, _i64p), _ptr(out, _i64p))
    return out


class Canonical:
    def __init__(self, T: np.ndarray, E: np.ndarray):
        self.T = T
        self.E = E

    @property
    def sha256(self) -> str:
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.T, dtype="<i4").tobytes())
        h.update(np.ascontiguousarray(self.E, dtype="<i4").tobytes())
        return h.hexdigest()

    def __eq__(self, other) -> bool:
        return (self.T.shape == other.T.shape and self.E.shape == other.E.shape
                and bool(np.array_equal(self.T, other.T)) and bool(np.array_equal(self.E, other.E)))


def _unique_rows(a: np.ndarray) -> np.ndarray:
    if a.shape[0] == 0:
        return a
    a = a[np.lexsort(a.T[::-1])]
    keep = np.ones(a.shape[0], bool)
    keep[1:] = np.any(a[1:] != a[:-1], axis=1)
    return a[keep]


def canonical_from_simplices(verts: np.ndarray, vertex_input: np.ndarray, dim: int,
                             ghost_ids: np.ndarray | None = None) -> Canonical:
    """verts: live simplices as vertex ids; vertex_input: vertex id -> input index (< 0 for
    super/ghost vertices).  Ghost simplices (any ghost vertex id) are dropped first."""
    verts = np.asarray(verts, np.int64)
    vin = np.asarray(vertex_input, np.int64)
    nsv = 2 * (dim + 1)
    ghost = np.any((verts >= dim + 1) & (verts < nsv), axis=1)
    verts = verts[~ghost]
    mapped = vin[verts]
    real = np.all(mapped >= 0, axis=1)
    T = np.sort(mapped[real], axis=1).astype(np.int32)
    T = _unique_rows(T)
    pairs = []
    M = dim + 1
    for i in range(M):
        for j in range(i + 1, M):
            a, b = mapped[:, i], mapped[:, j]
            ok = (a >= 0) & (b >= 0)
            pairs.append(np.stack([np.minimum(a[ok], b[ok]), np.maximum(a[ok], b[ok])], axis=1))
    E = _unique_rows(np.concatenate(pairs).astype(np.int32)) if pairs else np.zeros((0, 2), np.int32)
    return Canonical(T.reshape(-1, M), E.reshape(-1, 2))


def canonical_from_real(simplices_input: np.ndarray, dim: int) -> Canonical:
    """simplices already mapped to input indices with -1 for super vertices (no ghosts)."""
    s = np.asarray(simplices_input, np.int64)
    vin = np.arange(s.max() + 2 if s.size else 1, dtype=np.int64)
    # shift: -1 (super) -> map to a negative input through a sentinel id
    shifted = np.where(s < 0, vin.shape[0] - 1, s)
    vmap = vin.copy()
    vmap[-1] = -1
    M = dim + 1
    mapped = vmap[shifted]
    real = np.all(mapped >= 0, axis=1)
    T = _unique_rows(np.sort(mapped[real], axis=1).astype(np.int32)).reshape(-1, M)
    pairs = []
    for i in range(M):
        for j in range(i + 1, M):
            a, b = mapped[:, i], mapped[:, j]
            ok = (a >= 0) & (b >= 0)
            pairs.append(np.stack([np.minimum(a[ok], b[ok]), np.maximum(a[ok], b[ok])], axis=1))
    E = _unique_rows(np.concatenate(pairs).astype(np.int32)).reshape(-1, 2)
    return Canonical(T, E)
