"""Writes the golden fixtures under tests/golden/.  Run from the repository root:

    python tests/golden/make_golden.py

Nothing here runs the reference (it is Rust and cannot be built in this image).  The
fixtures pin the oracle from outside it:
  * qhull_*.json / qhull_small.npz : scipy.spatial.Delaunay (Qhull) run on P u S, where S is
    the reference's super simplex (src/delaunay_tree.rs:392-406, :550-558) -> canonical
    hashes and counts, and full canonical arrays for small sets;
  * brute_small.npz : exact-rational empty-circumsphere enumeration over all (d+1)-subsets
    of P u S for tiny P (independent of every floating-point routine);
  * kat.json : known answers -- the reference test's circumsphere KAT
    (tests/test_geometry.rs:5-15), max_simplex_id after new (tests/test_delaunay_tree.rs:20,44),
    the super-simplex trigonometric constants, and the C1 bounding sphere bits;
  * c1_points.npy : the 3D n=10,000 seed-0 uniform input (pins the generator too);
  * oracle_hashes.json : the oracle's exact-mode canonical hashes of inputs too slow for it to
    run inside a GPU test (a 32^3 block of the 1/128 lattice with 1e-9 jitter: about 30 s;
    200k clustered points); the oracle itself is pinned by the fixtures above.
    `python tests/golden/make_golden.py --oracle-only` rewrites just this file.
"""
from __future__ import annotations

import itertools
import json
import os
import struct
import sys
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402
from voronoids_amd import datasets  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden")


def hexd(x: float) -> str:
    return "0x%016x" % struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def qhull_canonical(P: np.ndarray) -> O.Canonical:
    from scipy.spatial import Delaunay
    dim = P.shape[1]
    S = O.super_vertices(P)[: dim + 1]
    D = Delaunay(np.vstack([S, P]))
    simp = D.simplices.astype(np.int64) - (dim + 1)
    simp[simp < 0] = -1
    return O.canonical_from_real(simp, dim)


def _det(m):
    n = len(m)
    if n == 1:
        return m[0][0]
    if n == 2:
        return m[0][0] * m[1][1] - m[0][1] * m[1][0]
    return sum(((-1) ** j) * m[0][j] * _det([row[:j] + row[j + 1:] for row in m[1:]]) for j in range(n))


def brute_canonical(P: np.ndarray) -> O.Canonical:
    """DT(P u S) by exact empty-sphere enumeration (general position assumed, checked)."""
    dim = P.shape[1]
    S = O.super_vertices(P)[: dim + 1]
    pts = [[Fraction(float(x)) for x in row] for row in np.vstack([S, P])]
    n = len(pts)
    simplices = []
    for comb in itertools.combinations(range(n), dim + 1):
        base = pts[comb[0]]
        o = _det([[pts[c][k] - base[k] for k in range(dim)] for c in comb[1:]])
        if o == 0:
            continue
        ok = True
        for q in range(n):
            if q in comb:
                continue
            rows = []
            for c in comb:
                r = [pts[c][k] - pts[q][k] for k in range(dim)]
                rows.append(r + [sum(x * x for x in r)])
            s = _det(rows)
            assert s != 0, "input not in general position"
            inside = (s * o > 0) if dim == 2 else (s * o < 0)
            if inside:
                ok = False
                break
        if ok:
            simplices.append([c - (dim + 1) if c > dim else -1 for c in comb])
    return O.canonical_from_real(np.array(simplices, np.int64), dim)


ORACLE_CASES = {
    "lattice_block_32_of_128": lambda: datasets.lattice_block(32, side=128, seed=7),
    "clustered_3d_200k": lambda: datasets.clustered(200_000, seed=3),
}


def oracle_hashes() -> None:
    out = {}
    for name, make in ORACLE_CASES.items():
        P = make()
        c = O.Tree.delaunay(P, O.MODE_EXACT).canonical()
        out[name] = {"n": int(P.shape[0]), "dim": int(P.shape[1]), "nT": int(c.T.shape[0]),
                     "nE": int(c.E.shape[0]), "sha256": c.sha256}
        print(name, out[name])
    with open(os.path.join(OUT, "oracle_hashes.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main() -> None:
    if "--oracle-only" in sys.argv:
        oracle_hashes()
        return
    summary = {}
    c1 = datasets.uniform(10_000, 3, seed=0)
    np.save(os.path.join(OUT, "c1_points.npy"), c1)
    for name, P in (("c1_3d_10k", c1), ("uniform_2d_10k", datasets.uniform(10_000, 2, seed=0)),
                    ("clustered_3d_20k", datasets.clustered(20_000, seed=0)),
                    ("uniform_3d_50k", datasets.uniform(50_000, 3, seed=1))):
        c = qhull_canonical(P)
        summary[name] = {"n": int(P.shape[0]), "dim": int(P.shape[1]), "nT": int(c.T.shape[0]),
                         "nE": int(c.E.shape[0]), "sha256": c.sha256}
        print(name, summary[name])
    with open(os.path.join(OUT, "qhull_hashes.json"), "w") as f:
        json.dump(summary, f, indent=1, sort_keys=True)

    small = {}
    for dim, n, seed in ((3, 300, 3), (2, 300, 4), (3, 60, 5), (2, 60, 6)):
        P = datasets.uniform(n, dim, seed=seed)
        c = qhull_canonical(P)
        key = f"d{dim}_n{n}_s{seed}"
        small[key + "_P"] = P
        small[key + "_T"] = c.T
        small[key + "_E"] = c.E
    np.savez_compressed(os.path.join(OUT, "qhull_small.npz"), **small)

    brute = {}
    for dim, n, seed in ((3, 9, 7), (2, 14, 8)):
        P = datasets.uniform(n, dim, seed=seed)
        c = brute_canonical(P)
        key = f"d{dim}_n{n}_s{seed}"
        brute[key + "_P"] = P
        brute[key + "_T"] = c.T
        brute[key + "_E"] = c.E
        print("brute", key, c.T.shape, c.E.shape)
    np.savez_compressed(os.path.join(OUT, "brute_small.npz"), **brute)

    cen, rad = O.bounding_sphere(c1)
    kat = {
        "circumsphere_corner_tet": {"verts": [[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]],
                                    "center": [0.5, 0.5, 0.5], "radius": 0.8660254037844386,
                                    "source": "tests/test_geometry.rs:5-15"},
        "max_simplex_id_after_new": {"3": 4, "2": 3, "source": "tests/test_delaunay_tree.rs:20,44"},
        "trig_hex": {"cos_2pi_3": "0xbfdffffffffffffc", "sin_2pi_3": "0x3febb67ae8584cab",
                     "cos_4pi_3": "0xbfe0000000000004", "sin_4pi_3": "0xbfebb67ae8584ca8"},
        "c1_bounding_sphere_hex": {"center": [hexd(x) for x in cen], "radius": hexd(rad)},
        "c1_super_vertices_hex": [[hexd(x) for x in row] for row in O.super_vertices(c1)],
        "reference_test_2d_points": [[0.3, 0.1], [1.0, 0.2], [0.1, 1.0], [0.5, 0.5]],
    }
    with open(os.path.join(OUT, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    oracle_hashes()


if __name__ == "__main__":
    main()
