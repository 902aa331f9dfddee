"""vo_predicates_noalloc (the allocation-free variant of vo_predicates: its staging and expansion
workspaces live in a static device array) against vo_predicates through the C ABI, on two calls in
a row -- the second smaller, so that it runs over what the first left in the array -- and the
size limit it rejects beyond."""
import ctypes
import zlib

import numpy as np
import pytest

from voronoids_amd import _lib
from voronoids_amd.batch import PRED_TYPES
from voronoids_amd.datasets import PREDICATE_FAMILIES as FAMILIES

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
MAX_N = 1 << 16


def call(fn, name, P):
    t, dim, arity = PRED_TYPES[name]
    p = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, arity, dim)
    n = p.shape[0]
    flat = p.reshape(-1, dim)
    c = np.zeros(3)
    c[:dim] = (flat.min(axis=0) + flat.max(axis=0)) / 2
    signs = np.full(n, 7, np.int32)
    tiers = np.full(n, 7, np.int32)
    rc = getattr(_lib.lib(), fn)(0, t, p.ctypes.data_as(_dp), n, c.ctypes.data_as(_dp), signs.ctypes.data_as(_i32p),
                                 tiers.ctypes.data_as(_i32p))
    return rc, signs, tiers


def mixed(name, n, seed):
    """instances from every family, so that all three tiers decide some of them"""
    parts = [gen(name, n // len(FAMILIES) + 1, seed=zlib.crc32(f"{name}/{fam}/{seed}".encode()) % 1000)
             for fam, gen in FAMILIES.items()]
    P = np.concatenate(parts)[:n]
    return P[np.random.default_rng(seed).permutation(P.shape[0])]


def test_rejects_more_than_the_maximum():
    _, dim, arity = PRED_TYPES["orient3d"]
    P = np.zeros((MAX_N + 1, arity, dim))
    rc, _, _ = call("vo_predicates_noalloc", "orient3d", P)
    assert rc == _lib.VO_ERR_INVALID
    assert b"exceed the maximum" in _lib.lib().vo_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["insphere", "orient3d", "incircle", "orient2d"])
def test_noalloc_equals_vo_predicates_twice(name):
    for n, seed in ((MAX_N, 1), (3001, 2)):
        P = mixed(name, n, seed)
        print(f"{name}: n={n} seed={seed}", flush=True)
        rc0, s0, t0 = call("vo_predicates", name, P)
        rc1, s1, t1 = call("vo_predicates_noalloc", name, P)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(s1, s0) and np.array_equal(t1, t0)
        if n == MAX_N:
            assert set(np.unique(t0).tolist()) == {1, 2, 3}
