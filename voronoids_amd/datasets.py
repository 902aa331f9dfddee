"""Deterministic synthetic point sets for the benchmark and the parity tests.

Every value comes from a counter-based SplitMix64 stream, so a set is a pure function of
(seed, stream, index) and any slice can be regenerated without generating the rest.
Uniform doubles are (x >> 11) * 2^-53, i.e. multiples of 2^-53 in [0, 1).
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def stream_key(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        k = _mix(np.array([np.uint64(seed) * _GOLDEN + np.uint64(stream)], dtype=np.uint64))
    return k[0]


def u64(seed: int, stream: int, start: int, count: int) -> np.ndarray:
    """values start .. start+count-1 of stream (seed, stream)"""
    key = stream_key(seed, stream)
    idx = np.arange(start, start + count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = key + (idx + np.uint64(1)) * _GOLDEN
    return _mix(z)


def unit(seed: int, stream: int, start: int, count: int) -> np.ndarray:
    return (u64(seed, stream, start, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform(n: int, dim: int, seed: int = 0, stream: int = 0) -> np.ndarray:
    """n points i.i.d. uniform in [0,1)^dim"""
    return unit(seed, stream, 0, n * dim).reshape(n, dim)


def clustered(n: int = 4_000_000, seed: int = 0, n_centers: int = 64, sigma: float = 0.02,
              frac_clustered: float = 0.8, dim: int = 3) -> np.ndarray:
    """frac_clustered of the points from isotropic Gaussians around n_centers uniform centers
    (component chosen uniformly, no clipping), the rest uniform in [0,1)^dim (3D by default)."""
    nc = int(round(n * frac_clustered))
    centers = uniform(n_centers, dim, seed, stream=11)
    comp = (u64(seed, 12, 0, nc) % np.uint64(n_centers)).astype(np.int64)
    u1 = unit(seed, 13, 0, nc * dim)
    u2 = unit(seed, 14, 0, nc * dim)
    g = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)    # Box-Muller, u1 in [0,1)
    pts_c = centers[comp] + sigma * g.reshape(nc, dim)
    pts_u = uniform(n - nc, dim, seed, stream=15)
    return np.ascontiguousarray(np.vstack([pts_c, pts_u]))


def lattice_jitter(side: int = 128, seed: int = 0, jitter: float = 1e-9) -> np.ndarray:
    """3D: (i/side, j/side, k/side) + i.i.d. uniform jitter in [-jitter, jitter) per coordinate"""
    g = np.arange(side, dtype=np.float64) / side
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    base = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    j = (2.0 * unit(seed, 16, 0, base.size) - 1.0) * jitter
    return np.ascontiguousarray(base + j.reshape(base.shape))


def lattice_block(block: int, side: int = 128, seed: int = 0, jitter: float = 1e-9) -> np.ndarray:
    """3D: the block^3 corner of the side^3 lattice, (i/side, j/side, k/side) for i, j, k < block,
    + i.i.d. uniform jitter in [-jitter, jitter) per coordinate (the lattice's own spacing)"""
    g = np.arange(block, dtype=np.float64) / side
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    base = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    j = (2.0 * unit(seed, 16, 0, base.size) - 1.0) * jitter
    return np.ascontiguousarray(base + j.reshape(base.shape))


def batch_sizes(n_sets: int = 4096, lo: int = 64, hi: int = 65536, seed: int = 0) -> np.ndarray:
    """set sizes floor(exp(U[ln lo, ln hi]))"""
    u = unit(seed, 17, 0, n_sets)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def batch_sets(sizes: np.ndarray, dim: int = 3, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """concatenated uniform sets and their offsets (len(sizes)+1)"""
    off = np.zeros(len(sizes) + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    return uniform(int(off[-1]), dim, seed, stream=18), off


# ------------------------------------------------------------------ predicate instances
# Tuples (n, arity, dim) of f64 points for one predicate, in families that make every
# precision tier decide some of them: generic, tight clusters, a jittered lattice, points
# rounded onto one circle / sphere, and exactly degenerate lattice tuples (sign 0).

PREDICATE_ARITY = {"orient2d": (3, 2), "incircle": (4, 2), "orient3d": (4, 3), "insphere": (5, 3)}


def tuples_uniform(name, n, seed):
    a, d = PREDICATE_ARITY[name]
    return unit(seed, 40, 0, n * a * d).reshape(n, a, d)


def tuples_clustered(name, n, seed, sigma=0.02):
    # tuples from one tight Gaussian cluster each (the clustered configuration's local geometry)
    a, d = PREDICATE_ARITY[name]
    rng = np.random.default_rng(seed)
    centers = rng.random((n, 1, d))
    return centers + sigma * rng.normal(size=(n, a, d))


def tuples_lattice_jitter(name, n, seed, side=128, jitter=1e-9):
    # neighboring points of a side^d lattice scaled to [0,1)^d, jittered by +-jitter
    a, d = PREDICATE_ARITY[name]
    rng = np.random.default_rng(seed)
    base = rng.integers(0, side - 2, size=(n, 1, d))
    offs = rng.integers(0, 2, size=(n, a, d))
    return (base + offs) / side + (2 * rng.random((n, a, d)) - 1) * jitter


def tuples_exact_degenerate(name, n, seed):
    # lattice points without jitter: many exactly coplanar / cospherical tuples (sign 0)
    a, d = PREDICATE_ARITY[name]
    rng = np.random.default_rng(seed)
    return rng.integers(0, 3, size=(n, a, d)) / 8.0


def tuples_rounded_sphere(name, n, seed):
    # points rounded onto one circle / sphere: in-sphere determinants at the f64 rounding level
    a, d = PREDICATE_ARITY[name]
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, a, d))
    return 0.5 + 0.3 * u / np.linalg.norm(u, axis=2, keepdims=True)


PREDICATE_FAMILIES = {"uniform": tuples_uniform, "clustered": tuples_clustered, "lattice_jitter": tuples_lattice_jitter,
                      "exact_degenerate": tuples_exact_degenerate, "rounded_sphere": tuples_rounded_sphere}
