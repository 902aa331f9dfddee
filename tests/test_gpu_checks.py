"""check_delaunay on the device shown to fail where it must, with the expected counts computed
independently (the oracle's exact predicates over the exported simplices, or the oracle's own
check_delaunay on the same tree); the super simplex through the product against hand-derived
bits; the near-degenerate and clustered configurations at their real spacing against the oracle.

Every corruption here stays in range: a vertex moves (its id and every index untouched) or a
simplex's vertex entries swap among themselves, so the checks read only live records."""
import json
import os

import numpy as np
import pytest

import voronoids_amd as vo
from oracle import oracle as O
from voronoids_amd import datasets

from test_closed_form import CASES, bits, expected_with_ghosts

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def local_violations(S, NB, X):
    """vo_tri_check's count restated with the oracle's exact predicates over an export: per live
    simplex, 1 if it is not positively oriented, and per face with a neighbor, 1 if the adjacency
    does not point back or if the neighbor's opposite vertex is strictly inside (the raw in-sphere
    sign of the simplex's own vertex order)"""
    dim = X.shape[1]
    pred_o = O.orient3d if dim == 3 else O.orient2d
    pred_i = O.insphere if dim == 3 else O.incircle
    bad = 0
    for i in range(S.shape[0]):
        P = X[S[i]]
        if pred_o(*P) <= 0:
            bad += 1
        for f in range(dim + 1):
            j = NB[i, f]
            if j < 0:
                continue
            back = np.nonzero(NB[j] == i)[0]
            if back.size != 1:
                bad += 1
                continue
            if pred_i(*P, X[S[j, back[0]]]) > 0:
                bad += 1
    return bad


def build_pair(dim, n, seed):
    P = datasets.uniform(n, dim, seed=seed)
    t = vo.delaunay(P, seed=seed)
    o = O.Tree.delaunay(P, O.MODE_EXACT)
    T, E = t.canonical()
    ref = o.canonical()
    assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)
    return P, t, o


def interior_vertex(P, t):
    """reference id of the input point closest to the center of the box"""
    i = int(np.argmin(((P - 0.5) ** 2).sum(1)))
    return 2 * t.M + i, i


@pytest.mark.parametrize("dim", [3, 2])
def test_moved_vertex_counts_match_the_oracle(dim):
    P, t, o = build_pair(dim, 1500, 30 + dim)
    assert t.violations() == 0 and t.check_pairs() == 0 and o.check_delaunay() == 0
    vid, i = interior_vertex(P, t)
    spacing = P.shape[0] ** (-1.0 / dim)
    q = P[i] + 0.45 * spacing * np.ones(dim) / np.sqrt(dim)
    t.set_vertex(vid, q)
    o.set_vertex(vid, q)
    S, NB = t.export()
    X = t.coordinates()
    assert np.array_equal(X[vid], q)
    expect = local_violations(S, NB, X)
    print(f"dim={dim} moved vertex {vid}: local violations expected {expect}", flush=True)
    assert expect > 0
    assert t.violations() == expect
    # the reference's own all-pairs count, against the oracle's check_delaunay on the same tree
    pairs = o.check_delaunay()
    assert pairs > 0
    assert t.check_pairs() == pairs


@pytest.mark.parametrize("dim", [3, 2])
def test_stale_records_are_reported(dim):
    P, t, o = build_pair(dim, 1200, 40 + dim)
    vid, i = interior_vertex(P, t)
    S, NB = t.export()
    X0 = t.coordinates()
    # a move that keeps the triangulation Delaunay under both the old and the new coordinates
    q = P[i] + 1e-7
    X1 = X0.copy()
    X1[vid] = q
    assert local_violations(S, NB, X0) == 0 and local_violations(S, NB, X1) == 0
    t.set_vertex(vid, q, records=False)
    holding = int((S == vid).any(axis=1).sum())
    assert holding >= dim + 1
    assert t.violations() == holding          # one record mismatch per simplex of the vertex
    t.set_vertex(vid, q, records=True)
    assert t.violations() == 0


@pytest.mark.parametrize("dim", [3, 2])
def test_swapped_vertices_are_reported(dim):
    P, t, o = build_pair(dim, 1000, 50 + dim)
    S, _ = t.export()
    # an all-real simplex in the middle of the export
    k = int(np.nonzero((S >= 2 * t.M).all(axis=1))[0][S.shape[0] // 3])
    t.swap_simplex_vertices(k, 0, 1)
    S2, NB2 = t.export()
    assert S2[k, 0] == S[k, 1] and S2[k, 1] == S[k, 0]
    expect = local_violations(S2, NB2, t.coordinates())
    assert expect >= 1
    assert t.violations() == expect
    t.swap_simplex_vertices(k, 0, 1)
    assert t.violations() == 0


@pytest.mark.parametrize("case", range(len(CASES)))
def test_product_super_vertices_hand_derived(case):
    P, _, _, sv = CASES[case]
    t = vo.DelaunayTree(P)
    X = t.coordinates()[: 2 * t.M]
    assert [[bits(x) for x in row] for row in X] == expected_with_ghosts(sv)
    t.add_points_to_tree(P)
    assert t.check_delaunay()


def oracle_hash(name):
    with open(os.path.join(G, "oracle_hashes.json")) as f:
        return json.load(f)[name]


def test_lattice_at_config_spacing_matches_oracle():
    # the near-degenerate configuration's own spacing (1/128) and jitter (1e-9): a 32^3 block;
    # the oracle's exact answer is cached (it takes the oracle about 30 s)
    ref = oracle_hash("lattice_block_32_of_128")
    P = datasets.lattice_block(32, side=128, seed=7)
    print(f"lattice block 32^3 at spacing 1/128, n={P.shape[0]}", flush=True)
    t = vo.delaunay(P, seed=5)
    T, E = t.canonical()
    assert (T.shape[0], E.shape[0]) == (ref["nT"], ref["nE"])
    assert O.Canonical(T, E).sha256 == ref["sha256"]
    assert t.check_delaunay()
    st = t.stats()
    assert st["tier2"][1] > 0                   # the FP64 filter was needed


def test_clustered_200k_matches_oracle():
    ref = oracle_hash("clustered_3d_200k")
    P = datasets.clustered(200_000, seed=3)
    print(f"clustered n={P.shape[0]}", flush=True)
    t = vo.delaunay(P, seed=1)
    T, E = t.canonical()
    assert (T.shape[0], E.shape[0]) == (ref["nT"], ref["nE"])
    assert O.Canonical(T, E).sha256 == ref["sha256"]
    assert t.check_delaunay()
