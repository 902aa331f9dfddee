"""The Python surface under the reference's module name: README.md:29-33 run verbatim, and the
PyDelauanyTree views (src/lib.rs:62-102) against the oracle's restatement of the reference's
maps -- super and ghost vertices, ghost simplices, super-incident simplices, adjacency and
incidence lists -- compared through the vertex tuple of each simplex (simplex ids are the build's
own)."""
import numpy as np
import pytest

from oracle import oracle as O
from voronoids_amd import datasets

pytestmark = pytest.mark.gpu

README_SNIPPET = """
import numpy as np
import voronoids

pts = np.random.uniform(size=(10000,3))
delaunay_graph = voronoids.delaunay(pts)
"""


def tuples_of(simplices: dict) -> dict:
    """simplex id -> sorted vertex tuple (unique per simplex: ghosts hold distinct ghost vertices)"""
    return {k: tuple(sorted(s.vertices)) for k, s in simplices.items()}


def test_readme_snippet_runs_verbatim():
    scope = {}
    exec(compile(README_SNIPPET, "README.md", "exec"), scope)
    g = scope["delaunay_graph"]
    pts = scope["pts"]
    assert type(g).__name__ == "PyDelauanyTree"
    sims, verts = g.simplices, g.vertices
    assert len(verts) == 8 + pts.shape[0]
    assert g.max_simplex_id >= max(sims)
    real = [s for s in sims.values() if min(s.vertices) >= 8]
    T = np.sort(np.array([s.vertices for s in real]) - 8, axis=1)
    ref = O.Tree.delaunay(pts, O.MODE_EXACT).canonical()
    assert np.array_equal(T[np.lexsort(T.T[::-1])], ref.T)
    for v in (0, 4, 8, 8 + pts.shape[0] - 1):
        assert all(v in sims[k].vertices for k in verts[v].simplex)


@pytest.mark.parametrize("dim,n", [(3, 2000), (2, 2000)])
def test_views_equal_the_oracle_maps(dim, n):
    import voronoids
    P = datasets.uniform(n, dim, seed=60 + dim)
    g = voronoids.delaunay(P)
    o = O.Tree.delaunay(P, O.MODE_EXACT)
    ex = o.export()
    M = dim + 1
    # the oracle's maps keyed by vertex tuple
    o_tup = {int(i): tuple(sorted(v)) for i, v in zip(ex["ids"], ex["verts"])}
    o_nb = {o_tup[int(i)]: sorted(o_tup[int(x)] for x in nb if x >= 0) for i, nb in zip(ex["ids"], ex["nbrs"])}
    sims = g.simplices
    tup = tuples_of(sims)
    assert sorted(tup.values()) == sorted(o_tup.values())            # ghosts and super-incident too
    for k, s in sims.items():
        assert sorted(tup[x] for x in s.neighbors) == o_nb[tup[k]], k
    # ghost simplices keep the reference's ids, vertex order, center and radius
    ghosts = [[4, 0, 1, 2], [5, 0, 2, 3], [6, 0, 3, 1], [7, 1, 2, 3]] if dim == 3 else [[3, 0, 1], [4, 0, 2], [5, 1, 2]]
    for k, gv in enumerate(ghosts):
        s = sims[k + 1]
        assert s.vertices == gv and s.center == [0.0] * dim and s.radius == 0.0 and len(s.neighbors) == 1
    # circumspheres by the reference's formulas (the vertex order within a simplex is the build's,
    # so the last bits may differ from the oracle's stored spheres)
    o_sph = {o_tup[int(i)]: (c, r) for i, c, r in zip(ex["ids"], ex["centers"], ex["radii"])}
    for k, s in list(sims.items())[:: 37]:
        if k <= M:
            continue
        c, r = o_sph[tup[k]]
        assert np.allclose(s.center, c, rtol=1e-9, atol=1e-12 * max(1.0, abs(r)))
        assert np.isclose(s.radius, r, rtol=1e-9)
    # vertices: coordinates bit for bit, incidence lists as tuple sets
    verts = g.vertices
    assert sorted(verts) == list(range(ex["coords"].shape[0]))
    for v in list(range(2 * M)) + list(range(2 * M, 2 * M + n, 97)):
        assert verts[v].point == ex["coords"][v].tolist(), v
        mine = sorted(tup[x] for x in verts[v].simplex)
        theirs = sorted(o_tup[x] for x in o.vertex_simplices(v))
        assert mine == theirs, v
    assert g.max_simplex_id >= max(sims)


def test_views_before_the_first_insertion():
    import voronoids
    P = datasets.uniform(50, 3, seed=1)
    t = voronoids.DelaunayTree(P)
    o = O.Tree(P)
    ex = o.export()
    assert t.max_simplex_id == o.max_simplex_id == 4
    sims = t.simplices
    assert sorted(sims) == [0, 1, 2, 3, 4]
    assert sorted(sims[0].vertices) == [0, 1, 2, 3] and sorted(sims[0].neighbors) == [1, 2, 3, 4]
    # simplex 0 keeps new()'s stored sphere: (bounding-sphere center, 10 x radius)
    i0 = int(np.nonzero(ex["ids"] == 0)[0][0])
    assert sims[0].center == ex["centers"][i0].tolist() and sims[0].radius == ex["radii"][i0]
    verts = t.vertices
    for v in range(8):
        assert sorted(verts[v].simplex) == sorted(o.vertex_simplices(v)), v


def _live_tuples(tree_export):
    return {tuple(sorted(int(x) for x in v)) for v in tree_export["verts"]}


@pytest.mark.parametrize("dim", [3, 2])
def test_locate_is_the_cavity_of_the_next_insertion(dim):
    # locate + get_new_simplices (src/delaunay_tree.rs:33-123): the simplices the insertion of q
    # kills and the ones it creates, read off the oracle's trees before and after inserting q
    import voronoids_amd as vo
    n = 3000
    P = datasets.uniform(n, dim, seed=70 + dim)
    t = vo.delaunay(P, seed=2)
    before = t.canonical()
    S, NB = t.export()
    rng = np.random.default_rng(dim)
    for trial in range(3):
        q = rng.random(dim)
        killed, facets = t.locate(q)
        A = _live_tuples(O.Tree.delaunay(P).export())
        B = _live_tuples(O.Tree.delaunay(np.vstack([P, q])).export())
        qid = 2 * (dim + 1) + n
        assert list(killed) == sorted(killed)
        assert sorted(tuple(sorted(S[k])) for k in killed) == sorted(A - B)
        new = {tuple(sorted([qid] + row[2:].tolist())) for row in facets}
        assert new == {s for s in B if qid in s}
        for row in facets:
            assert row[0] in set(killed.tolist())
            assert (row[1] == -1) == (int(row[1]) not in range(S.shape[0]))
            if row[1] >= 0:
                assert row[1] not in set(killed.tolist()) and row[0] in NB[row[1]]
    # a query changes nothing; a vertex is refused as the reference's locate refuses it
    T2, E2 = t.canonical()
    assert np.array_equal(T2, before[0]) and np.array_equal(E2, before[1])
    with pytest.raises(vo.VoError) as e:
        t.locate(P[17])
    assert e.value.code == -4


def test_locate_before_the_first_insertion():
    import voronoids_amd as vo
    P = datasets.uniform(100, 3, seed=3)
    t = vo.DelaunayTree(P)
    killed, facets = t.locate(P[0])
    assert killed.tolist() == [0] and facets.shape == (4, 5) and (facets[:, 1] == -1).all()
    assert sorted(sorted(r[2:].tolist()) for r in facets) == [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]
