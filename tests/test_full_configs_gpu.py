"""The measured configurations at their full sizes, through the C ABI, with the size-independent
checks each one allows (the oracle cannot run at these sizes):

* 3D clustered n = 4,000,000 (64 Gaussian clusters, sigma 0.02) and the 128^3 lattice with 1e-9
  jitter at the configuration's spacing 1/128: the exact device check (every face locally
  Delaunay, adjacency and orientation consistent), the vertex count, the super simplex's four
  hull faces, and the same simplex set from a second insertion order (the Delaunay
  triangulation of points in general position is unique, so any order must give it);
* batching, 4,096 3D sets with sizes log-uniform in [64, 65,536] (37.7M points): every set
  completes and every 16th set's simplex set equals the one-set path's on the same points;
* serving, requests of 1e3..1e6 points in 2D and 3D submitted while others run: each result
  equals the one-set path's."""
import time

import numpy as np
import pytest

import voronoids_amd as vo
from voronoids_amd import batch as B
from voronoids_amd import datasets

pytestmark = pytest.mark.gpu


def _full_check(P, seeds=(0, 1)):
    n = P.shape[0]
    hashes = []
    for seed in seeds:
        t = vo.DelaunayTree(P)
        t.add_points_to_tree(P, seed=seed)
        assert t.violations() == 0
        nv, ns, _ = t.counts()
        st = t.stats()
        assert nv == n + 4 and st["points_inserted"] == n and st["points_duplicate"] == 0
        s, nb = t.export()
        assert np.count_nonzero(nb < 0) == 4
        hashes.append(B.simplex_set_hash(s))
        tiers = st
        t.close()
    assert hashes[0] == hashes[1]
    return tiers


def test_clustered_4m():
    st = _full_check(datasets.clustered(4_000_000, seed=0))
    # clusters push some in-sphere tests past the FP32 tier
    assert st["tier2"][1] > 0


def test_lattice_128_cubed_with_jitter():
    st = _full_check(datasets.lattice_jitter(128, seed=0))
    # 1e-9 jitter on a 1/128 lattice: near-cospherical cells need the FP64 filter and the
    # expansions
    assert st["tier2"][1] > 0 and st["exact_evaluated"] > 0


def test_batch_4096_sets():
    sizes = datasets.batch_sizes(4096, 64, 65536, seed=0)
    P, off = datasets.batch_sets(sizes, 3, seed=0)
    eng = B.Engine(3, max_points=18_000_000, max_simplices=(1 << 29) - 4096)
    try:
        res = eng.batch_flat(P, off, 0)
    finally:
        eng.close()
    assert len(res) == 4096 and all(r.status == 0 for r in res)
    assert [r.n_points for r in res] == sizes.tolist()
    plain = B.PlainPath()
    try:
        for k in range(0, 4096, 16):
            s = plain.simplices(P[off[k]:off[k + 1]], k)[0]
            assert B.simplex_set_hash(res[k].simplices) == B.simplex_set_hash(s), k
    finally:
        plain.close()


def test_served_requests_up_to_a_million_points():
    rng = np.random.default_rng(90)
    reqs = []
    for k in range(40):
        dim = 2 if k % 2 else 3
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
        reqs.append((dim, datasets.uniform(n, dim, seed=91, stream=k), 2000 + k))
    srv = vo.Server(max_points_2d=1 << 22, max_points_3d=1 << 22)
    got = {}
    try:
        for k, (dim, P, seed) in enumerate(reqs):
            srv.submit(P, seed=seed, tag=k)
            for r in srv.results():
                got[r.tag] = r
        t0 = time.time()
        while len(got) < len(reqs) and time.time() - t0 < 120:
            for r in srv.results(timeout_ms=200):
                got[r.tag] = r
    finally:
        srv.close()
    assert len(got) == len(reqs)
    plain = B.PlainPath()
    try:
        for k, (dim, P, seed) in enumerate(reqs):
            r = got[k]
            assert r.dim == dim and r.status == 0 and r.n_points == P.shape[0]
            s = plain.simplices(P, seed)[0]
            assert B.simplex_set_hash(r.simplices) == B.simplex_set_hash(s), k
    finally:
        plain.close()
