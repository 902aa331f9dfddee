"""Precision tiers predicate by predicate (vo_predicates through the C ABI) against the oracle's
exact predicates (big-integer determinants, a different method from the GPU's expansions):
zero sign disagreements on generic, clustered, jittered-lattice, rounded-sphere and exactly
degenerate tuples, and every tier decides some of them."""
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from voronoids_amd import batch as B
from voronoids_amd.datasets import PREDICATE_FAMILIES as FAMILIES

pytestmark = pytest.mark.gpu

ORACLE = {"orient2d": O.orient2d, "incircle": O.incircle, "orient3d": O.orient3d, "insphere": O.insphere}


@pytest.mark.parametrize("name", ["orient2d", "incircle", "orient3d", "insphere"])
def test_tiers_agree_with_the_exact_oracle(name):
    tiers_seen = set()
    for fam, gen in FAMILIES.items():
        P = gen(name, 3000, seed=zlib.crc32(f"{name}/{fam}".encode()) % 1000)
        signs, tiers = B.predicates(name, P)
        ref = np.array([ORACLE[name](*P[i]) for i in range(P.shape[0])])
        bad = np.nonzero(signs != ref)[0]
        assert bad.size == 0, (fam, bad[:5], signs[bad[:5]], ref[bad[:5]], tiers[bad[:5]])
        tiers_seen |= set(np.unique(tiers).tolist())
        if fam == "exact_degenerate":
            # a zero comes from the FP64 filter only where its arithmetic is provably exact
            # (Shewchuk's stage A), else from the expansions; never from the FP32 tier
            assert (ref == 0).any() and (tiers[ref == 0] >= 2).all()
    assert tiers_seen == {1, 2, 3}, tiers_seen
