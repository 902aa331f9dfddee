"""GPU construction (libvoronoids_hip.so through the C ABI) against the oracle, the golden
fixtures and the size-independent invariants of a Delaunay triangulation."""
import json
import os

import numpy as np
import pytest

import voronoids_amd as vo
from oracle import oracle as O
from voronoids_amd import datasets

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


def canon_sha(T, E):
    return O.Canonical(T, E).sha256


def test_max_simplex_id_after_new():
    # tests/test_delaunay_tree.rs:20 and :44
    assert vo.DelaunayTree(datasets.uniform(1000, 3, seed=0)).max_simplex_id == 4
    assert vo.DelaunayTree(np.array([[0.3, 0.1], [1.0, 0.2], [0.1, 1.0], [0.5, 0.5]])).max_simplex_id == 3


@pytest.mark.parametrize("name", ["c1_3d_10k", "uniform_2d_10k", "clustered_3d_20k", "uniform_3d_50k"])
def test_golden_hashes(name):
    with open(os.path.join(G, "qhull_hashes.json")) as f:
        ref = json.load(f)[name]
    P = {"c1_3d_10k": lambda: np.load(os.path.join(G, "c1_points.npy")),
         "uniform_2d_10k": lambda: datasets.uniform(10_000, 2, seed=0),
         "clustered_3d_20k": lambda: datasets.clustered(20_000, seed=0),
         "uniform_3d_50k": lambda: datasets.uniform(50_000, 3, seed=1)}[name]()
    print(f"case {name} n={P.shape[0]}", flush=True)
    t = vo.delaunay(P, seed=0)
    T, E = t.canonical()
    assert (T.shape[0], E.shape[0]) == (ref["nT"], ref["nE"])
    assert canon_sha(T, E) == ref["sha256"]


def test_small_fixtures_and_brute_force():
    for fn in ("qhull_small.npz", "brute_small.npz"):
        z = np.load(os.path.join(G, fn))
        for k in sorted({k.rsplit("_", 1)[0] for k in z.files}):
            t = vo.delaunay(z[k + "_P"], seed=3)
            T, E = t.canonical()
            assert np.array_equal(T, z[k + "_T"]), (fn, k)
            assert np.array_equal(E, z[k + "_E"]), (fn, k)


@pytest.mark.parametrize("dim,n,seed", [(3, 3000, 21), (2, 3000, 22), (3, 30000, 23), (2, 100000, 24)])
def test_matches_oracle(dim, n, seed):
    P = datasets.uniform(n, dim, seed=seed)
    t = vo.delaunay(P, seed=seed)
    T, E = t.canonical()
    ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
    assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)


def test_insertion_order_independent():
    P = datasets.uniform(20000, 3, seed=31)
    a = vo.delaunay(P, seed=1).canonical()
    b = vo.delaunay(P, seed=99).canonical()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_reference_test_sequence_3d():
    # tests/test_delaunay_tree.rs:7-38: new(vertices), insert 100 of them, then 1000 more
    V = datasets.uniform(1000, 3, seed=40)
    V2 = datasets.uniform(1000, 3, seed=41)
    t = vo.DelaunayTree(V)
    t.add_points_to_tree(V[:100], seed=5)
    t.add_points_to_tree(V2, seed=6)
    assert t.check_delaunay()
    o = O.Tree(V)
    o.insert_seq(V[:100])
    o.add_points_parallel(V2, input_base=100, nthreads=2)
    ref = o.canonical()
    T, E = t.canonical()
    assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)


def test_invariants_2d_count_and_check():
    n = 300_000
    t = vo.delaunay(datasets.uniform(n, 2, seed=50))
    nv, ns, _ = t.counts()
    assert nv == n + 3 and ns == 2 * n + 1          # DT(P u S): 2(n+3) - 2 - 3 triangles
    assert t.violations() == 0


def test_invariants_3d_euler_and_check():
    n = 200_000
    t = vo.delaunay(datasets.uniform(n, 3, seed=51))
    s, nb = t.export()
    assert t.violations() == 0
    V = n + 4
    T = s.shape[0]
    faces = set()
    edges = set()
    for a in range(4):
        for b in range(a + 1, 4):
            e = np.sort(s[:, [a, b]], axis=1)
            edges.update(map(tuple, e))
    F = (4 * T + np.count_nonzero(nb < 0)) // 2
    assert V - len(edges) + F - T == 1               # Euler characteristic of a 3-ball
    assert np.count_nonzero(nb < 0) == 4             # the super simplex hull


def test_duplicates_raise_like_the_reference():
    P = datasets.uniform(5000, 3, seed=60)
    P[100] = P[7]
    with pytest.raises(vo.VoError) as e:
        vo.delaunay(P)
    assert e.value.code == -4


def test_near_degenerate_lattice_matches_oracle():
    # 12^3 lattice with 1e-9 jitter: FP32 cannot separate it, FP64 and the exact tier must
    g = (np.arange(12) / 12.0)
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    P = X + (2 * datasets.unit(0, 70, 0, X.size).reshape(X.shape) - 1) * 1e-9
    t = vo.delaunay(P, seed=2)
    T, E = t.canonical()
    ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
    assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)
    st = t.stats()
    assert st["tier2"][1] > 0


def test_exact_tier_decides_what_fp64_cannot():
    # points rounded onto a sphere / circle: in-sphere determinants sit at the f64 rounding
    # level, the FP64 filter fails and the expansion tier decides (no exact degeneracy:
    # the rounding errors are generic, so the triangulation is still unique)
    rng = np.random.default_rng(72)
    solo = 0
    for dim in (3, 2):
        u = rng.normal(size=(400, dim))
        S = 0.5 + 0.3 * u / np.linalg.norm(u, axis=1)[:, None]
        P = np.vstack([S, 0.5 + 0.05 * rng.random((50, dim))])
        t = vo.delaunay(P, seed=4)
        T, E = t.canonical()
        ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
        assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E), dim
        st = t.stats()
        assert st["exact_evaluated"] > 0, dim
        solo += st["solo_runs"]
        assert t.violations() == 0
    assert solo > 0          # some cavities outgrew the per-point LDS lists and went solo


@pytest.mark.parametrize("dim", [2, 3])
def test_exact_tier_alone_builds_the_same_triangulation(dim):
    # filters off: every predicate is queued, evaluated by the expansion kernel, cached, and
    # read back from the cache on the point's next attempt
    P = datasets.uniform(300, dim, seed=80 + dim)
    t = vo.DelaunayTree(P)
    t.config("filters", 0)
    t.add_points_to_tree(P, seed=1)
    T, E = t.canonical()
    ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
    assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)
    st = t.stats()
    assert st["tier1"] == [0, 0] and st["tier2"] == [0, 0]
    assert st["tier3"][0] > 0 and st["tier3"][1] > 0 and st["exact_evaluated"] > 0
    assert t.violations() == 0


@pytest.mark.parametrize("lanes", [1, 4, 8, 16, 32])
@pytest.mark.parametrize("dim", [2, 3])
def test_every_lane_group_size_matches_oracle(dim, lanes):
    P = datasets.uniform(20000, dim, seed=90 + dim)
    t = vo.DelaunayTree(P)
    t.config("group_lanes", lanes)
    t.add_points_to_tree(P, seed=7)
    T, E = t.canonical()
    ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
    assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)
    assert t.violations() == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_compaction_keeps_the_triangulation(dim):
    # compact after every few rounds: pending points' conflict pointers, walk hints and
    # adjacency are renumbered many times; output, check, counts and later inserts stay exact
    P = datasets.uniform(30000, dim, seed=110 + dim)
    Q = datasets.uniform(5000, dim, seed=120 + dim) * 0.5 + 0.25
    t = vo.DelaunayTree(np.vstack([P, Q]))
    t.config("reuse", 0)                      # append-only records: compaction has dead ones to drop
    t.config("compact_min", 256)
    t.config("compact_ratio_milli", 1500)
    t.add_points_to_tree(P, seed=5)
    t.add_points_to_tree(Q, seed=6)
    st = t.stats()
    assert st["compactions"] >= 3
    T, E = t.canonical()
    ref = O.Tree.delaunay(np.vstack([P, Q]), O.MODE_EXACT).canonical()
    assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)
    assert t.violations() == 0
    nv, ns, mx = t.counts()
    assert ns == (t.export()[0].shape[0]) and mx == st["created"] + dim + 1 - 1 + 1


@pytest.mark.parametrize("dim", [2, 3])
def test_record_reuse_keeps_records_near_the_live_count(dim):
    # new simplices take their winner's killed records: records stay close to the live simplices
    # (append-only: about 4x (3D) / 3x (2D) as many), with the same triangulation either way
    P = datasets.uniform(100_000, dim, seed=130 + dim)
    out = {}
    for reuse in (1, 0):
        t = vo.DelaunayTree(P)
        t.config("reuse", reuse)
        t.add_points_to_tree(P, seed=7)
        st = t.stats()
        out[reuse] = (t.canonical(), st["records"], st["created"], t.counts())
        assert t.violations() == 0
    (T1, E1), rec1, cr1, (nv1, ns1, mx1) = out[1]
    (T0, E0), rec0, cr0, (nv0, ns0, mx0) = out[0]
    assert np.array_equal(T1, T0) and np.array_equal(E1, E0) and ns1 == ns0
    assert rec1 <= 1.25 * ns1 + 64 and rec0 >= 2.5 * ns0
    assert mx1 == cr1 + dim + 1 and mx0 == cr0 + dim + 1
    ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
    assert np.array_equal(T1, ref.T) and np.array_equal(E1, ref.E)


@pytest.mark.parametrize("dim", [2, 3])
def test_xcd_region_windows_keep_the_triangulation(dim):
    # windows sorted into 8 spatial regions, block b serving region b % 8, with frequent
    # compactions: same triangulation as the oracle's
    P = datasets.uniform(60000, dim, seed=130 + dim)
    t = vo.DelaunayTree(P)
    t.config("xcd_regions", 1)
    t.config("reuse", 0)                      # append-only records: compaction has dead ones to drop
    t.config("compact_ratio_milli", 1500)
    t.config("compact_min", 256)
    t.add_points_to_tree(P, seed=3)
    T, E = t.canonical()
    ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
    assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)
    assert t.violations() == 0 and t.stats()["compactions"] > 0


@pytest.mark.parametrize("dim", [2, 3])
def test_device_planned_rounds_match_host_planned(dim):
    # the device planner applies the host's window rule: the same rounds and windows and the
    # same triangulation for any number of rounds per chunk, and a second insert continues
    # from the planner's state
    P = datasets.uniform(120000, dim, seed=140 + dim)
    Q = datasets.uniform(20000, dim, seed=150 + dim) * 0.5 + 0.25
    ref = O.Tree.delaunay(np.vstack([P, Q]), O.MODE_EXACT).canonical()
    seen = []
    # boot: the first rounds in one workgroup (k_boot) off, or on up to 256-point windows;
    # flags: the commit's per-slot outcomes compacted by the planning launch, or appended
    for pipelined, chunk, boot, flags in ((0, 4, 0, 1), (1, 1, 0, 1), (1, 4, 0, 1), (1, 7, 0, 1), (1, 4, 256, 1),
                                          (1, 1, 128, 1), (1, 4, 0, 0)):
        t = vo.DelaunayTree(np.vstack([P, Q]))
        t.config("outcome_flags", flags)
        t.config("abort_early", 0)              # the schedule that does not depend on timing
        t.config("walk_budget", 0)              # (a walk budget pauses walks whose length depends on
        t.config("pipelined", pipelined)        # which concurrent write a hint cell kept)
        t.config("chunk_rounds", chunk)
        t.config("boot_window", boot)
        t.add_points_to_tree(P, seed=8)
        t.add_points_to_tree(Q, seed=9)
        T, E = t.canonical()
        assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E), (pipelined, chunk, boot, flags)
        st = t.stats()
        # (rounds run inside the one-workgroup kernel are not launches of the round kernels)
        seen.append((st["rounds"], st["points_inserted"], st["windows_total"], st["created"], st["killed"],
                     st["launches_locate_cavity"] if boot == 0 else None))
        assert t.violations() == 0
    assert all(x[:5] == seen[0][:5] for x in seen), seen
    assert all(x[5] == seen[0][5] for x in seen if x[5] is not None), seen


@pytest.mark.parametrize("dim", [3, 2])
def test_allocation_chunks_and_outcome_flags_equal_the_oracle(dim):
    # simplex records from one counter or from the planner's per-round chunks, and the next
    # window's points appended through one counter or written per slot and compacted by the
    # planning launch: the exact triangulation either way; the chunks' unused tails stay
    # fresh records (never live, never exported), so records exceed the created simplices only
    # with the chunks on
    P = datasets.uniform(60000, dim, seed=170 + dim)
    ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
    for stripes, flags in ((0, 0), (0, 1), (1, 0), (1, 1)):
        print(f"case dim={dim} alloc_stripes={stripes} outcome_flags={flags}", flush=True)
        t = vo.DelaunayTree(P)
        t.config("alloc_stripes", stripes)
        t.config("outcome_flags", flags)
        t.add_points_to_tree(P, seed=12)
        T, E = t.canonical()
        assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E), (dim, stripes, flags)
        st = t.stats()
        assert t.violations() == 0
        # records: the root, then one per created simplex, plus the chunks' unused tails (each
        # planned round reserves 64 chunks of at least 16 records)
        extra = st["records"] - st["created"] - 1
        assert (extra > 64 * 16) if stripes else (0 <= extra < 1024), (dim, stripes, flags, st["records"], st["created"])
    t = vo.DelaunayTree(P[:100])
    for key, bad in (("alloc_stripes", 2), ("outcome_flags", -1), ("boot_window", 70000), ("no_such_knob", 1)):
        with pytest.raises(vo.VoError):
            t.config(key, bad)


def test_device_planned_rounds_hand_over_exact_and_solo_rounds():
    # points rounded onto a sphere / circle: rounds queue exact predicates and some cavities
    # outgrow the per-point LDS lists (solo runs); the planner halts on each such round and the
    # host finishes it, with the host loop's rounds, windows and result
    rng = np.random.default_rng(160)
    solo = 0
    for dim in (3, 2):
        u = rng.normal(size=(500, dim))
        P = np.vstack([0.5 + 0.3 * u / np.linalg.norm(u, axis=1)[:, None], 0.5 + 0.05 * rng.random((60, dim))])
        ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
        seen = []
        # the one-workgroup first rounds (boot 256) halt for the host on such rounds as well, and
        # so do rounds whose pending list is appended rather than compacted (flags 0)
        for pipelined, boot, flags in ((0, 0, 1), (1, 0, 1), (1, 256, 1), (1, 0, 0)):
            t = vo.DelaunayTree(P)
            t.config("outcome_flags", flags)
            t.config("abort_early", 0)
            t.config("walk_budget", 0)
            t.config("pipelined", pipelined)
            t.config("boot_window", boot)
            t.add_points_to_tree(P, seed=11)
            T, E = t.canonical()
            assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E), (dim, pipelined, boot, flags)
            st = t.stats()
            assert st["exact_evaluated"] > 0
            seen.append((st["rounds"], st["points_inserted"], st["windows_total"], st["solo_runs"]))
        assert all(x == seen[0] for x in seen), (dim, seen)
        solo += seen[1][3]
    assert solo > 0


def test_full_size_3d_4m_invariants():
    # the headline workload at full size (3D uniform n = 4,000,000): the exact device check
    # (local Delaunay property of every face, adjacency and orientation consistency), vertex
    # count, the super simplex's four hull faces, and (without early abort and the walk budget,
    # whose winners depend on timing) the same rounds and windows from the device planner as from
    # the host loop
    n = 4_000_000
    P = datasets.uniform(n, 3, seed=0)
    seen = []
    for pipelined, abort in ((1, 1), (1, 0), (0, 0)):
        t = vo.DelaunayTree(P)
        t.config("abort_early", abort)
        if not abort:
            t.config("walk_budget", 0)
        t.config("pipelined", pipelined)
        t.add_points_to_tree(P, seed=0)
        assert t.violations() == 0
        nv, ns, mx = t.counts()
        st = t.stats()
        assert nv == n + 4 and st["points_inserted"] == n and st["points_duplicate"] == 0
        assert mx == st["created"] + 4          # d + 1 after DelaunayTree::new, plus one per created simplex
        if not abort:
            seen.append((st["rounds"], st["windows_total"], st["created"], st["killed"], ns))
        if pipelined:
            _, nb = t.export()
            assert np.count_nonzero(nb < 0) == 4
        t.close()
    assert seen[0] == seen[1], seen


def test_full_size_2d_10m_invariants():
    # 2D uniform n = 10,000,000: exact device check and the triangle count of DT(P u S)
    n = 10_000_000
    t = vo.delaunay(datasets.uniform(n, 2, seed=0))
    nv, ns, _ = t.counts()
    assert nv == n + 3 and ns == 2 * n + 1
    assert t.violations() == 0
    t.close()


def test_full_size_3d_4m_order_and_engine_independent():
    # DT(P u S) is unique for points in general position, so at the headline size two insertion
    # orders (seeds) and the multi-set engine must give the same simplex set (order-free hash of
    # the sorted vertex tuples in the reference id space)
    from voronoids_amd.batch import simplex_set_hash
    P = datasets.uniform(4_000_000, 3, seed=0)
    hashes, canon = [], []
    for seed in (0, 1):
        t = vo.delaunay(P, seed=seed)
        s, _ = t.export()
        hashes.append(simplex_set_hash(s))
        canon.append(t.canonical())          # sorted on the device
        t.close()
    r = vo.batch_triangulate([P], seed=2)[0]
    hashes.append(simplex_set_hash(r.simplices))
    assert hashes[0] == hashes[1] == hashes[2], hashes
    assert np.array_equal(canon[0][0], canon[1][0]) and np.array_equal(canon[0][1], canon[1][1])
    assert canon[0][0].shape[1] == 4 and canon[0][1].shape[1] == 2


@pytest.mark.parametrize("dim,n", [(3, 1), (3, 5), (2, 1), (2, 7), (3, 250_000), (2, 600_000)])
def test_device_canonical_equals_host_canonical(dim, n):
    # the canonical form is sorted on the device (radix sorts); the host routine the engine's
    # results use must give the same arrays, from tiny sets (no real simplex) to many tiles
    from voronoids_amd.batch import canonical_from_simplices
    P = datasets.uniform(n, dim, seed=170 + n % 97)
    t = vo.delaunay(P, seed=3)
    T, E = t.canonical()
    s, _ = t.export()
    Th, Eh = canonical_from_simplices(s, dim)
    assert T.shape == Th.shape and E.shape == Eh.shape
    assert np.array_equal(T, Th) and np.array_equal(E, Eh)
    if n > 1000:
        assert T.shape[0] > 0 and E.shape[0] > 0
        assert np.all(np.diff(E[:, 0]) >= 0)
    t.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [3, 2])
def test_abort_early_same_triangulation(dim):
    # early abort changes which points win a round, never the result: uniform points and points
    # rounded onto a sphere (exact predicates, solo runs) give the oracle's triangulation with
    # abort_early 0 and 1, and the engine agrees with both
    rng = np.random.default_rng(170 + dim)
    u = rng.normal(size=(3000, dim))
    S = 0.5 + 0.3 * u / np.linalg.norm(u, axis=1)[:, None]
    for P in (datasets.uniform(150000, dim, seed=180 + dim), np.vstack([S, 0.5 + 0.05 * rng.random((200, dim))])):
        ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
        for abort in (0, 1):
            t = vo.DelaunayTree(P)
            t.config("abort_early", abort)
            t.add_points_to_tree(P, seed=3)
            T, E = t.canonical()
            assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E), (dim, abort, P.shape)
            assert t.violations() == 0
            t.close()
        parts = [P[:40000], P[40000:]] if P.shape[0] > 40000 else [P]
        refs = [O.Tree.delaunay(part, O.MODE_EXACT).canonical() for part in parts]
        for abort in (0, 1):
            res = vo.batch_triangulate(parts, seed=4, abort_early=abort)
            for ref_k, r in zip(refs, res):
                T_k, E_k = r.canonical()
                assert np.array_equal(T_k, ref_k.T) and np.array_equal(E_k, ref_k.E), (dim, abort, "engine")


_TRANSFORMS = {
    "offset1e6": lambda U: U + 1e6,
    "offset1e9": lambda U: U + 1e9,
    "neg_offset": lambda U: U - 123456.789,
    "scale1e-9": lambda U: U * 1e-9,
    "scale1e-6": lambda U: U * 1e-6,
    "scale1e6": lambda U: U * 1e6,
    "scale1e12": lambda U: U * 1e12,
    "aniso": lambda U: U * np.array([1.0, 1e-4, 1e-4][:U.shape[1]]),
}


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("name", sorted(_TRANSFORMS))
def test_translated_and_scaled_inputs_match_oracle(dim, name):
    # the FP32 tier works on coordinates translated to the bounding sphere's center and the filters'
    # error bounds scale with the input: far offsets, tiny and huge scales and a 1e4:1 box give the
    # exact oracle's triangulation (the 3D box sends thousands of predicates to the exact tier)
    P = _TRANSFORMS[name](datasets.uniform(20000, dim, seed=300 + dim))
    t = vo.delaunay(P, seed=1)
    T, E = t.canonical()
    ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
    assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E), (dim, name)
    assert t.violations() == 0
    t.close()


@pytest.mark.parametrize("dim", [3, 2])
def test_reset_handle_equals_a_fresh_tree(dim):
    # vo_tri_reset: one handle, many sets in a row (larger, smaller, tiny, shifted), each equal to
    # a fresh tree's triangulation and the oracle's; the round keys go on decreasing across sets
    from voronoids_amd import batch as B
    plain = B.PlainPath()
    try:
        for k, (n, shift) in enumerate(((20000, 0.0), (3000, 5.0), (50, -2.0), (40000, 0.5))):
            P = datasets.uniform(n, dim, seed=140 + k) + shift
            s, rc = plain.simplices(P, seed=k)
            s0, rc0 = B.plain_simplices(P, seed=k)
            assert rc == 0 and rc0 == 0
            assert B.simplex_set_hash(s) == B.simplex_set_hash(s0), (k, n)
            T, E = B.canonical_from_simplices(s, dim)
            ref = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
            assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E), (k, n)
    finally:
        plain.close()
    t = vo.DelaunayTree(datasets.uniform(1000, dim, seed=1))
    t.add_points_to_tree(datasets.uniform(1000, dim, seed=1))
    r1 = t.stats()["rounds"]
    P2 = datasets.uniform(500, dim, seed=2)
    t.reset(P2)
    assert t.counts()[:2] == (dim + 1, 1) and t.max_simplex_id == dim + 1
    t.add_points_to_tree(P2)
    assert 0 < t.stats()["rounds"] <= r1 and t.check_delaunay()
    assert t.simplices[dim + 2].vertices is not None and len(t.vertices) == 2 * (dim + 1) + 500
