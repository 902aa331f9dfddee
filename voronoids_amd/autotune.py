"""Autotuner for the round kernels: sweeps lanes per window point (fixed sizes and the
per-round rule at several lane targets) and the initial window ratio on a workload, times
each setting with the library's hipEvent clock, and writes the best setting per dimension to
voronoids_amd/tuning.json, which DelaunayTree applies when it exists.

    python -m voronoids_amd.autotune --dim 3 --n 1000000
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING = os.path.join(HERE, "tuning.json")


def candidates(dim: int) -> list[dict]:
    out = [{"group_lanes": g} for g in (1, 4, 8, 16, 32)]
    for lt in (1 << 16, 1 << 17, 1 << 18, 1 << 19, 1 << 20, 1 << 21):
        out.append({"group_lanes": 0, "lanes_target": lt})
    for gm in (100, 400):
        out.append({"group_lanes": 0, "window_gamma_milli": gm})
    return out


def time_setting(P: np.ndarray, setting: dict, reps: int) -> float:
    import voronoids_amd as vo
    best = float("inf")
    for _ in range(reps):
        t = vo.DelaunayTree(P)
        for k, v in setting.items():
            t.config(k, v)
        t.stage(P, seed=0)
        t.run()
        best = min(best, t.stats()["ms_run"])
        t.close()
    return best


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--dim", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--write", action="store_true", help="store the best setting in tuning.json")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    from voronoids_amd import datasets
    P = datasets.uniform(args.n, args.dim, seed=0)
    results = []
    for s in candidates(args.dim):
        ms = time_setting(P, s, args.reps)
        results.append((ms, s))
        print(json.dumps({"ms": round(ms, 3), **s}), flush=True)
    results.sort(key=lambda r: r[0])
    best_ms, best = results[0]
    print(json.dumps({"best": best, "ms": best_ms, "dim": args.dim, "n": args.n}), flush=True)
    if args.write:
        table = {}
        if os.path.exists(TUNING):
            with open(TUNING) as f:
                table = json.load(f)
        table[str(args.dim)] = {"setting": best, "ms": best_ms, "n": args.n}
        with open(TUNING, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
