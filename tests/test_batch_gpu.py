"""Batched and served modes (vo_engine / vo_batch_triangulate / vo_server through the C ABI):
every set's canonical output equals the one-set path's on the same points and seed, which the
one-set tests pin to the oracle and the golden fixtures; a few sets are also checked against
the oracle directly.  The engine paths that only small arenas or odd inputs reach (range
extension, admission waits, window scaling, solo cavities, duplicates, empty sets, reuse of
freed arena space) are forced here with small capacities."""
import time

import numpy as np
import pytest

import voronoids_amd as vo
from oracle import oracle as O
from voronoids_amd import datasets

pytestmark = pytest.mark.gpu


def log_uniform_sizes(n_sets, lo, hi, seed):
    u = datasets.unit(seed, 5, 0, n_sets)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def make_sets(sizes, dim, seed):
    return [datasets.uniform(int(n), dim, seed=seed, stream=100 + k) for k, n in enumerate(sizes)]


def plain_canonical(P, seed):
    t = vo.DelaunayTree(P)
    try:
        t.add_points_to_tree(P, seed=seed)
        return t.canonical()
    finally:
        t.close()


def assert_same_as_plain(sets, results, seed):
    assert len(results) == len(sets)
    for k, (P, r) in enumerate(zip(sets, results)):
        assert r.tag == k and r.status == 0 and r.n_points == P.shape[0], (k, r.status)
        T, E = r.canonical()
        Tp, Ep = plain_canonical(P, seed + k)
        assert np.array_equal(T, Tp) and np.array_equal(E, Ep), f"set {k} n={P.shape[0]}"


@pytest.mark.parametrize("dim", [3, 2])
def test_batch_matches_plain_and_oracle(dim):
    sizes = log_uniform_sizes(48, 64, 6000, seed=10 + dim)
    sets = make_sets(sizes, dim, seed=11)
    print(f"batch dim={dim} sets={len(sets)} points={int(sizes.sum())}", flush=True)
    eng = vo.Engine(dim, max_points=1 << 20)
    res = eng.batch(sets, seed=5)
    st = eng.stats()
    eng.close()
    assert st["sets_done"] == len(sets) and st["replays"] > 0
    assert_same_as_plain(sets, res, 5)
    for k in (0, int(np.argmax(sizes))):
        ref = O.Tree.delaunay(sets[k], O.MODE_EXACT).canonical()
        T, E = res[k].canonical()
        assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)


def test_batch_small_arena_waits_and_extends():
    # vertex arena for a few sets at a time, first simplex ranges of 3 per point: sets queue
    # for admission and every set needs further ranges
    sizes = log_uniform_sizes(40, 64, 1500, seed=20)
    sets = make_sets(sizes, 3, seed=21)
    eng = vo.Engine(3, max_points=6000, max_simplices=200_000, tets_per_point_milli=3000)
    res = eng.batch(sets, seed=9)
    st = eng.stats()
    eng.close()
    print(st, flush=True)
    assert st["admission_waits"] > 0 and st["range_extensions"] > 0
    assert_same_as_plain(sets, res, 9)


def test_batch_window_cap_and_group_sizes():
    sizes = log_uniform_sizes(24, 200, 5000, seed=30)
    sets = make_sets(sizes, 3, seed=31)
    # (window_cap 256 carries pending points over a round while the commit's outcome flags are
    # appended after them; outcome_flags 0 appends through the list's counter instead)
    for cfg in ({"window_cap": 256}, {"group_lanes": 1}, {"group_lanes": 8, "rounds_per_replay": 2},
                {"group_lanes": 32, "min_window": 64}, {"outcome_flags": 0}, {"outcome_flags": 0, "window_cap": 256}):
        eng = vo.Engine(3, max_points=1 << 18, **cfg)
        res = eng.batch(sets, seed=13)
        eng.close()
        print("config", cfg, flush=True)
        assert_same_as_plain(sets, res, 13)


def test_batch_near_cospherical_sets_take_the_solo_path():
    # points rounded onto a sphere / circle: in-sphere signs at the f64 rounding level (the
    # exact tier decides) and cavities that outgrow the lane-group lists (solo runs)
    rng = np.random.default_rng(40)
    solo = 0
    for dim, lanes in ((3, 32), (2, 16)):
        sets = []
        for k in range(6):
            u = rng.normal(size=(400, dim))
            S = 0.5 + 0.3 * u / np.linalg.norm(u, axis=1)[:, None]
            sets.append(np.vstack([S, 0.5 + 0.05 * rng.random((50, dim))]))
        sets += make_sets([500, 2000, 64], dim, seed=41)
        eng = vo.Engine(dim, max_points=1 << 16, group_lanes=lanes)
        res = eng.batch(sets, seed=4)
        st = eng.stats()
        eng.close()
        print(dim, st, flush=True)
        solo += st["solo_runs"]
        assert_same_as_plain(sets, res, 4)
        for k in range(2):
            ref = O.Tree.delaunay(sets[k], O.MODE_EXACT).canonical()
            T, E = res[k].canonical()
            assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)
    assert solo > 0


def test_batch_exact_tier_only():
    sets = make_sets([300, 150, 64, 500], 2, seed=45)
    eng = vo.Engine(2, max_points=1 << 16, filters=0)
    res = eng.batch(sets, seed=2)
    eng.close()
    assert_same_as_plain(sets, res, 2)


def test_batch_duplicates_empty_and_tiny_sets():
    sets = make_sets([0, 1, 2, 3, 4, 5, 17, 300], 3, seed=50)
    dup = datasets.uniform(1000, 3, seed=51)
    dup[10] = dup[3]
    dup[500] = dup[3]
    sets.append(dup)
    eng = vo.Engine(3, max_points=1 << 16)
    res = eng.batch(sets, seed=1)
    eng.close()
    assert res[-1].status == -4 and res[-1].points_duplicate == 2
    assert [r.status for r in res[:-1]] == [0] * (len(sets) - 1)
    for k, (P, r) in enumerate(zip(sets[:-1], res[:-1])):
        T, E = r.canonical()
        Tp, Ep = plain_canonical(P, 1 + k)
        assert np.array_equal(T, Tp) and np.array_equal(E, Ep), k
    # the duplicate set: the one-set path drops the same two points
    t = vo.DelaunayTree(dup)
    with pytest.raises(vo.VoError):
        t.add_points_to_tree(dup, seed=1 + len(sets) - 1)
    T, E = res[-1].canonical()
    Tp, Ep = t.canonical()
    t.close()
    assert np.array_equal(T, Tp) and np.array_equal(E, Ep)


def test_engine_reuses_freed_space_across_batches():
    eng = vo.Engine(3, max_points=20000, max_simplices=700_000)
    for b in range(3):
        sets = make_sets(log_uniform_sizes(12, 64, 4000, seed=60 + b), 3, seed=61 + b)
        res = eng.batch(sets, seed=100 * b)
        assert_same_as_plain(sets, res, 100 * b)
    st = eng.stats()
    eng.close()
    assert st["sets_done"] == 36 and st["sets_resident"] == 0 and st["simplex_in_use"] == 0


def test_server_results_equal_alone_and_in_the_mix():
    srv = vo.Server(max_points_2d=1 << 20, max_points_3d=1 << 20)
    reqs = []
    rng = np.random.default_rng(70)
    for k in range(24):
        dim = 2 if k % 2 else 3
        n = int(np.exp(rng.uniform(np.log(1000), np.log(30000))))
        reqs.append((dim, datasets.uniform(n, dim, seed=71, stream=k), 1000 + k))
    got = {}
    for k, (dim, P, seed) in enumerate(reqs):
        srv.submit(P, seed=seed, tag=k)
        time.sleep(0.002 * (k % 3))
        for r in srv.results():
            got[r.tag] = r
    t0 = time.time()
    while len(got) < len(reqs) and time.time() - t0 < 60:
        for r in srv.results(timeout_ms=200):
            got[r.tag] = r
    srv.close()
    assert len(got) == len(reqs)
    for k, (dim, P, seed) in enumerate(reqs):
        r = got[k]
        assert r.dim == dim and r.status == 0
        assert r.t_submit <= r.t_admit <= r.t_done
        T, E = r.canonical()
        Tp, Ep = plain_canonical(P, seed)
        assert np.array_equal(T, Tp) and np.array_equal(E, Ep), k
