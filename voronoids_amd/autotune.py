"""Autotuner of the round kernels (k_cavity_g / k_commit_g): build constants and runtime knobs.

Two kinds of choice, per dimension:
  build    compile-time constants of voronoids_amd/csrc/kernels_group.hpp: the least waves per
           SIMD each round kernel is compiled for, which sets its register budget
           (VO_CAVITY_WAVES / VO_COMMIT_WAVES in 3D, VO_CAVITY_WAVES2 / VO_COMMIT_WAVES2 in 2D),
           and the LDS list capacities per window point -- cavity entries and boundary facets
           (VO_GROUP_KC3 / VO_GROUP_FC3, VO_GROUP_KC2 / VO_GROUP_FC2; a point whose cavity
           outgrows them is inserted alone) -- with the bytes of facet tuples kept for ridge
           matching (VO_MATCH_BYTES3, and VO_MATCH_MIN3: the least tuples per point, which gives
           the small lane groups more).  Each candidate is its own build of the library
           (voronoids_amd/variants/), timed in its own process.
  runtime  vo_tri_config knobs of that build: lanes per window point (a fixed size, or the
           per-round rule at another lane target, for both round kernels or the commit alone),
           the window's initial ratio, the winner fractions the large and the small windows aim
           at, the least window, rounds per submitted chunk, and the walk's step budget per round.

The search: every build candidate at the default runtime setting, then every runtime candidate on
the fastest build, then the union of the runtime settings that each gained on their own.  The
winner must also not lose on clustered points (64 Gaussian clusters): it is timed there against
the defaults and rejected, the next one tried, when slower by more than the margin.  A candidate counts only if its triangulation has no violation and as many
simplices as the default's.  Times are medians of vo_tri_run's device time; the winner replaces
the defaults only when it is more than 1% faster (run-to-run noise is about that).

    python -m voronoids_amd.autotune build                      # compile the build candidates (hipcc)
    python -m voronoids_amd.autotune sweep --dim 3 --n 4000000 --write
    python -m voronoids_amd.autotune sweep --dim 2 --n 10000000 --write

--write stores the winner per dimension in voronoids_amd/tuning.json: voronoids_amd/build.py
compiles its build constants into the shipped library, and every DelaunayTree applies its runtime
setting.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING = os.path.join(HERE, "tuning.json")
VARIANTS = os.path.join(HERE, "variants")

# build constants and their defaults (kernels_group.hpp); a candidate changes one group of them
BUILD_DEFAULTS = {
    3: {"VO_CAVITY_WAVES": 3, "VO_COMMIT_WAVES": 4, "VO_GROUP_KC3": 192, "VO_GROUP_FC3": 256, "VO_MATCH_BYTES3": 4096,
        "VO_MATCH_MIN3": 0},
    2: {"VO_CAVITY_WAVES2": 3, "VO_COMMIT_WAVES2": 4, "VO_GROUP_KC2": 64, "VO_GROUP_FC2": 96},
}
BUILD_CANDIDATES = {
    3: [{}, {"VO_CAVITY_WAVES": 2}, {"VO_CAVITY_WAVES": 4}, {"VO_COMMIT_WAVES": 3}, {"VO_COMMIT_WAVES": 5},
        {"VO_GROUP_KC3": 160, "VO_GROUP_FC3": 224}, {"VO_GROUP_KC3": 256, "VO_GROUP_FC3": 320},
        {"VO_MATCH_BYTES3": 2048}, {"VO_MATCH_BYTES3": 8192}, {"VO_MATCH_MIN3": 64}],
    2: [{}, {"VO_CAVITY_WAVES2": 2}, {"VO_CAVITY_WAVES2": 4}, {"VO_COMMIT_WAVES2": 3}, {"VO_COMMIT_WAVES2": 5},
        {"VO_GROUP_KC2": 48, "VO_GROUP_FC2": 80}, {"VO_GROUP_KC2": 96, "VO_GROUP_FC2": 128}],
}


def runtime_candidates(dim: int) -> list[dict]:
    out: list[dict] = [{}]
    base_lt = 1 << 21 if dim == 3 else 1 << 16
    for lt in (base_lt >> 2, base_lt >> 1, base_lt << 1, base_lt << 2):
        out.append({"lanes_target": lt})
    out += [{"group_lanes": g} for g in ((4, 8, 16) if dim == 3 else (4, 8))]
    for gm in (100, 200, 400):
        out.append({"window_gamma_milli": gm})
    lo = 650 if dim == 3 else 550
    for d in (-100, -50, 50):
        out.append({"big_frac_lo_milli": lo + d})
    out += [{"chunk_rounds": 2}, {"chunk_rounds": 8}]
    for lt in (base_lt >> 1, base_lt << 1, base_lt << 2):
        out.append({"commit_lanes_target": lt})
    out += [{"small_frac_lo_milli": 250, "small_frac_hi_milli": 500}, {"small_frac_lo_milli": 200, "small_frac_hi_milli": 450},
            {"small_frac_lo_milli": 450, "small_frac_hi_milli": 700}]
    out += [{"min_window": 128}, {"min_window": 256}]
    out += [{"walk_budget": 16}, {"walk_budget": 32}, {"walk_budget": 64}]
    boot = 0 if dim == 3 else 256
    out += [{"boot_window": b} for b in (0, 128, 256, 512) if b != boot]
    return out


def tag_of(defines: dict) -> str:
    return "default" if not defines else "_".join(f"{k[3:].lower()}-{v}" for k, v in sorted(defines.items()))


def variant_path(dim: int, defines: dict) -> str:
    return os.path.join(VARIANTS, f"libvoronoids_hip.d{dim}.{tag_of(defines)}.so")


def tuned_build(dim: int) -> dict:
    """build constants of the stored winner for one dimension ({} when none is stored)"""
    if not os.path.exists(TUNING):
        return {}
    with open(TUNING) as f:
        return dict(json.load(f).get(str(dim), {}).get("build", {}))


def tuned_defines() -> tuple:
    """"NAME=VALUE" for every stored build constant of both dimensions (the macros are disjoint)"""
    out = []
    for dim in (3, 2):
        for k, v in sorted(tuned_build(dim).items()):
            if k not in BUILD_DEFAULTS[dim]:
                raise ValueError(f"tuning.json: {k} is not a {dim}D build constant")
            out.append(f"{k}={int(v)}")
    return tuple(out)


def build_variants(dims=(3, 2), jobs: int = 4) -> list[str]:
    """compile every build candidate (each with the other dimension's stored constants)"""
    from concurrent.futures import ThreadPoolExecutor
    from .build import build
    os.makedirs(VARIANTS, exist_ok=True)
    todo = []
    for dim in dims:
        other = 2 if dim == 3 else 3
        for cand in BUILD_CANDIDATES[dim]:
            defs = tuple(f"{k}={v}" for k, v in sorted({**tuned_build(other), **cand}.items()))
            todo.append((variant_path(dim, cand), defs))

    def one(item):
        path, defs = item
        t0 = time.time()
        build(out=path, defines=defs, static=False, tuned=False)
        print(json.dumps({"built": os.path.relpath(path, HERE), "defines": list(defs),
                          "s": round(time.time() - t0, 1)}), flush=True)
        return path
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(one, todo))


def trial(dim: int, n: int, reps: int, settings: list[dict], workload: str = "uniform") -> list[dict]:
    """time vo_tri_run for each runtime setting in this process's library build (VORONOIDS_LIB),
    on uniform points or on clustered ones (64 Gaussian clusters, sigma 0.02, 80% of the points)"""
    sys.path.insert(0, os.path.dirname(HERE))
    import numpy as np
    import voronoids_amd as vo
    from voronoids_amd import datasets
    P = datasets.uniform(n, dim, seed=0) if workload == "uniform" else datasets.clustered(n, seed=0, dim=dim)
    out = []
    for s in settings:
        ms, sims, viol, err = [], None, None, None
        try:
            for r in range(reps):
                t = vo.DelaunayTree(P, tuned=False)
                for k, v in s.items():
                    t.config(k, v)
                t.stage(P, seed=0)
                t.run()
                ms.append(t.stats()["ms_run"])
                if r == 0:
                    sims = t.counts()[1]
                    viol = t.violations()
                t.close()
        except vo.VoError as e:
            err = str(e)
        rec = {"setting": s, "ms": float(np.median(ms)) if ms and err is None else None, "runs": ms,
               "simplices": sims, "violations": viol, "error": err}
        print(json.dumps(rec), flush=True)
        out.append(rec)
    return out


def _trial_in_process(lib: str, dim: int, n: int, reps: int, settings: list[dict], timeout: int,
                      workload: str = "uniform") -> list[dict]:
    env = dict(os.environ, VORONOIDS_LIB=lib)
    cmd = [sys.executable, "-u", "-m", "voronoids_amd.autotune", "trial", "--dim", str(dim), "--n", str(n),
           "--reps", str(reps), "--settings", json.dumps(settings), "--workload", workload]
    # the child's lines are echoed as they come, so a long sweep shows its progress
    out = []
    t0 = time.time()
    with subprocess.Popen(cmd, env=env, cwd=os.path.dirname(HERE), stdout=subprocess.PIPE, text=True) as proc:
        for line in proc.stdout:
            if line.startswith("{"):
                out.append(json.loads(line))
                print(f"# {os.path.basename(lib)}: {line.strip()}", file=sys.stderr, flush=True)
            if time.time() - t0 > timeout:
                proc.kill()
                raise RuntimeError(f"trial with {os.path.basename(lib)} passed {timeout} s")
        rc = proc.wait()
    if rc != 0:
        raise RuntimeError(f"trial with {os.path.basename(lib)} failed with exit status {rc}")
    return out


def pick(candidates: list[dict], margin: float) -> dict:
    """the fastest valid candidate that no check rejected, if it beats the default build at the
    default setting (the first candidate) by more than `margin` (a fraction: run-to-run noise),
    else that default"""
    default = candidates[0]
    best = min((c for c in candidates if c["valid"] and not c.get("rejected")), key=lambda c: c["ms"])
    return best if best["ms"] < default["ms"] * (1.0 - margin) else default


def sweep(dim: int, n: int, reps: int, write: bool, timeout: int = 300, margin: float = 0.01) -> dict:
    results = []

    def ok(rec: dict, ref) -> bool:
        return rec["ms"] is not None and rec["violations"] == 0 and rec["simplices"] == ref
    # 1. every build at the default runtime setting
    ref_sims = None
    builds = []
    for cand in BUILD_CANDIDATES[dim]:
        lib = variant_path(dim, cand)
        if not os.path.exists(lib):
            raise FileNotFoundError(f"{lib}: run `python -m voronoids_amd.autotune build` first")
        rec = _trial_in_process(lib, dim, n, reps, [{}], timeout)[0]
        if ref_sims is None:
            ref_sims = rec["simplices"]
        rec.update(build=cand, valid=ok(rec, ref_sims))
        print(json.dumps({"dim": dim, "build": tag_of(cand), "ms": rec["ms"], "valid": rec["valid"]}), flush=True)
        results.append(rec)
        if rec["valid"]:
            builds.append((rec["ms"], cand))
    default_ms = results[0]["ms"]
    best_build = min(builds, key=lambda b: b[0])[1]
    # 2. every runtime setting on the fastest build
    stage2 = []
    for rec in _trial_in_process(variant_path(dim, best_build), dim, n, reps, runtime_candidates(dim), timeout):
        rec.update(build=best_build, valid=ok(rec, ref_sims))
        print(json.dumps({"dim": dim, "build": tag_of(best_build), "setting": rec["setting"], "ms": rec["ms"],
                          "valid": rec["valid"]}), flush=True)
        results.append(rec)
        stage2.append(rec)
    # 3. the union of the runtime settings that each beat that build's default by more than the
    # margin (those that touch different knobs), tried once more together
    base_ms = stage2[0]["ms"]
    gains = sorted((r for r in stage2[1:] if r["valid"] and r["ms"] < base_ms * (1.0 - margin)), key=lambda r: r["ms"])
    merged: dict = {}
    for r in gains:
        if not set(r["setting"]) & set(merged):
            merged.update(r["setting"])
    if len(gains) > 1 and merged != gains[0]["setting"]:
        for rec in _trial_in_process(variant_path(dim, best_build), dim, n, reps, [merged], timeout):
            rec.update(build=best_build, valid=ok(rec, ref_sims))
            print(json.dumps({"dim": dim, "build": tag_of(best_build), "setting": rec["setting"], "ms": rec["ms"],
                              "valid": rec["valid"]}), flush=True)
            results.append(rec)
    candidates = [{"build": r["build"], "setting": r["setting"], "ms": r["ms"], "valid": r["valid"]} for r in results]
    # 4. the winner must not lose on clustered points: tried there against the defaults, and
    # rejected (the next one tried) when slower by more than the margin
    base = None
    while True:
        best = pick(candidates, margin)
        if best is candidates[0]:
            break
        if base is None:
            base = _trial_in_process(variant_path(dim, {}), dim, n, reps, [{}], timeout, "clustered")[0]
        rec = _trial_in_process(variant_path(dim, best["build"]), dim, n, reps, [best["setting"]], timeout,
                                "clustered")[0]
        best["clustered_ms"], best["clustered_default_ms"] = rec["ms"], base["ms"]
        if rec["ms"] is not None and rec["violations"] == 0 and rec["ms"] <= base["ms"] * (1.0 + margin):
            break
        best["rejected"] = "slower on clustered points"
        print(json.dumps({"dim": dim, "rejected": {k: best[k] for k in ("build", "setting", "ms", "clustered_ms",
                                                                       "clustered_default_ms")}}), flush=True)
    summary = {"build": best["build"], "setting": best["setting"], "ms": best["ms"], "default_ms": default_ms,
               "margin": margin, "n": n, "reps": reps, "workload": f"{dim}D uniform n={n}, seed 0",
               "candidates": candidates}
    print(json.dumps({"dim": dim, "best": {k: summary[k] for k in ("build", "setting", "ms", "default_ms")}}), flush=True)
    if write:
        table = {}
        if os.path.exists(TUNING):
            with open(TUNING) as f:
                table = json.load(f)
        table[str(dim)] = summary
        with open(TUNING, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
            f.write("\n")
    return summary


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=("build", "sweep", "trial"))
    ap.add_argument("--dim", type=int, default=3)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--settings", default="[{}]", help="trial: JSON list of runtime settings")
    ap.add_argument("--write", action="store_true", help="sweep: store the winner in tuning.json")
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--workload", choices=("uniform", "clustered"), default="uniform", help="trial: the points")
    args = ap.parse_args()
    n = args.n or (4_000_000 if args.dim == 3 else 10_000_000)
    if args.cmd == "build":
        build_variants(jobs=args.jobs)
    elif args.cmd == "trial":
        trial(args.dim, n, args.reps, json.loads(args.settings), args.workload)
    else:
        sweep(args.dim, n, args.reps, args.write)
    return 0


if __name__ == "__main__":
    sys.exit(main())
