"""Benchmark: inserted points per second of the MI355X Delaunay construction.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 3d4m|2d10m|clustered|lattice]

A step is one complete triangulation of one synthetic point set that is already resident in
HBM (created and staged outside the timed region); each timed run is bracketed by a device
synchronization on both sides.  The default workload is the headline configuration: 3D,
n = 4,000,000 points i.i.d. uniform in [0,1)^3 (float64, seed 0), random insertion order.

Prints one JSON line with the throughput, the roofline of the dominant kernel (algorithmic
bytes from counted events, priced with the per-event bytes of the HBM model below) and the
CPU baseline (the oracle's restatement of the reference's own CPU path, timed on this host).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# bytes per event (3D, 2D): insert = point + conflict pointer; killed simplex read (tv+ta);
# created simplex write; boundary facet (outside neighbor opposite vertex + coordinates +
# adjacency patch); distinct cavity vertex coordinates; relocation / walk step
EVENT_BYTES = {3: dict(insert=28, killed=32, created=32, facet=32, vertex=24, reloc=44),
               2: dict(insert=20, killed=24, created=24, facet=24, vertex=16, reloc=32)}

WORKLOADS = {
    "3d4m": dict(dim=3, n=4_000_000, kind="uniform",
                 label="3D Delaunay, n=4,000,000 i.i.d. uniform [0,1)^3 f64, seed 0, random insertion order"),
    "2d10m": dict(dim=2, n=10_000_000, kind="uniform",
                  label="2D Delaunay, n=10,000,000 i.i.d. uniform [0,1)^2 f64, seed 0, random insertion order"),
    "clustered": dict(dim=3, n=4_000_000, kind="clustered",
                      label="3D Delaunay, n=4,000,000: 80% from 64 Gaussians (sigma 0.02) + 20% uniform, f64"),
    "lattice": dict(dim=3, n=128 ** 3, kind="lattice",
                    label="3D Delaunay, 128^3 lattice / 128 with uniform jitter +-1e-9, f64"),
    "3d1m": dict(dim=3, n=1_000_000, kind="uniform",
                 label="3D Delaunay, n=1,000,000 i.i.d. uniform [0,1)^3 f64, seed 0"),
}


TRAFFIC_PROFILES = {"3d4m": "traffic_round_kernels_3d4m_uniform.md"}


def make_points(w: dict) -> np.ndarray:
    from voronoids_amd import datasets
    if w["kind"] == "uniform":
        return datasets.uniform(w["n"], w["dim"], seed=0)
    if w["kind"] == "clustered":
        return datasets.clustered(w["n"], seed=0)
    return datasets.lattice_jitter(128, seed=0)


def algorithmic_bytes(dim: int, n: int, st: dict) -> dict:
    """counted events x per-event bytes; relocations use min(counted walk steps, model)"""
    b = EVENT_BYTES[dim]
    ins, killed, created = st["points_inserted"], st["killed"], st["created"]
    verts = created / 2 + 2 * ins if dim == 3 else created
    reloc_model = n * (dim + 1) * max(math.log(max(n, 2)) - 0.42, 0.0)
    reloc = min(st["walk_steps"], reloc_model)
    cavity = ins * b["insert"] + killed * b["killed"] + created * b["facet"] + verts * b["vertex"] + reloc * b["reloc"]
    commit = created * b["created"]
    return {"k_cavity": cavity, "k_commit": commit, "relocations_counted": st["walk_steps"],
            "relocations_model": reloc_model}


def cpu_baseline(dim: int, threads: int) -> dict:
    """The oracle's restatement of voronoids.delaunay (src/lib.rs:104-125) with the
    reference's own arithmetic: 1e5 sequential inserts, then add_points_to_tree (make_queue
    / find_placement / level batches) in parallel on `threads` OpenMP threads."""
    from oracle import oracle as O
    from voronoids_amd import datasets
    n = 600_000 if dim == 3 else 2_000_000
    P = datasets.uniform(n, dim, seed=0)
    t0 = time.perf_counter()
    tr = O.Tree.delaunay(P, O.MODE_REF, nthreads=threads)
    dt = time.perf_counter() - t0
    del tr
    return {"value": n / dt, "unit": "pts/s", "cores": threads, "kind": "port",
            "sample": f"{dim}D uniform n={n} (1e5 sequential insert_point + {n - 100000} via add_points_to_tree), "
                      f"{dt:.1f} s wall"}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary of this
    workload (FETCH_SIZE and WRITE_SIZE passes, each scaled by what the counter reads for a
    device-to-device copy of known size), or None when no summary exists."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_PROFILES.get(workload, "-"))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_json: "):
            rec = json.loads(line[len("traffic_json: "):])
            k = rec["kernels"].get(kernel)
            return None if k is None else {"bytes_per_launch": k["traffic_per_launch"],
                                           "fetch_factor": rec["fetch_factor"], "write_factor": rec["write_factor"],
                                           "source": "profiles/" + TRAFFIC_PROFILES[workload]}
    return None


ENGINE_STAT_KEYS = ("rounds", "replays", "solo_runs", "range_extensions", "admission_waits", "max_sets_resident",
                    "ms_replays", "ms_boundary", "ms_admit", "ms_collect")


def run_batched(seed: int = 0, plain_stride: int = 8) -> dict:
    """Batching: 4,096 independent 3D sets, sizes floor(exp(U[ln 64, ln 65536])), points i.i.d.
    uniform in [0,1)^3, through one engine (vo_batch_triangulate: device-planned rounds
    replayed as HIP graphs, sets admitted by arena footprint as others complete).  Timed from
    host points to every set's simplices on the host.  The plain path (vo_tri create + insert +
    export + destroy, one set at a time) is timed on every plain_stride-th set, and each of
    those sets' simplex sets is compared with the batched result (order-free hash)."""
    from voronoids_amd import batch as B
    from voronoids_amd import datasets
    sizes = datasets.batch_sizes(4096, 64, 65536, seed=seed)
    P, off = datasets.batch_sets(sizes, 3, seed=seed)
    eng = B.Engine(3, max_points=18_000_000, max_simplices=(1 << 29) - 4096)
    try:
        eng.batch_flat(P[: off[64]], off[:65], seed)          # warm-up: graph capture, arena touch
        base = eng.stats()
        t0 = time.perf_counter()
        res = eng.batch_flat(P, off, seed)
        dt = time.perf_counter() - t0
        st = eng.stats()
    finally:
        eng.close()
    print(f"# batched: {len(res)} sets, {int(off[-1])} points in {dt:.2f} s", file=sys.stderr, flush=True)
    idx = list(range(0, len(sizes), plain_stride))
    t0 = time.perf_counter()
    hashes = []
    for k in idx:
        s, _ = B.plain_simplices(P[off[k]:off[k + 1]], seed + k)
        hashes.append(B.simplex_set_hash(s))
    dtp = time.perf_counter() - t0                            # includes the hashing: subtracted below
    t_hash = time.perf_counter()
    for k in idx:
        B.simplex_set_hash(res[k].simplices)
    dth = time.perf_counter() - t_hash
    dtp = max(dtp - dth, 1e-9)
    same = sum(B.simplex_set_hash(res[k].simplices) == h for k, h in zip(idx, hashes))
    pts_s = int(sum(sizes[k] for k in idx))
    print(f"# plain path: {len(idx)} sets, {pts_s} points in {dtp:.2f} s; identical {same}/{len(idx)}",
          file=sys.stderr, flush=True)
    ok = all(r.status == 0 for r in res)
    return {
        "workload": "4,096 3D sets, sizes log-uniform [64, 65536], i.i.d. uniform [0,1)^3 f64",
        "sets": len(res), "points": int(off[-1]), "s": dt, "sets_per_s": len(res) / dt, "pts_per_s": int(off[-1]) / dt,
        "plain": {"sets_timed": len(idx), "points_timed": pts_s, "s": dtp, "sets_per_s": len(idx) / dtp,
                  "pts_per_s": pts_s / dtp, "sample": f"every {plain_stride}th set"},
        "speedup_sets_per_s": (len(res) / dt) / (len(idx) / dtp),
        "speedup_pts_per_s": (int(off[-1]) / dt) / (pts_s / dtp),
        "identical_output": {"compared_sets": len(idx), "identical": int(same), "all_status_ok": ok},
        "engine": {k: st[k] - (base[k] if k not in ("max_sets_resident",) else 0) for k in ENGINE_STAT_KEYS},
    }


def _percentiles(x) -> dict:
    x = np.asarray(x, dtype=np.float64)
    return {"p50_ms": float(np.percentile(x, 50) * 1e3), "p99_ms": float(np.percentile(x, 99) * 1e3),
            "mean_ms": float(x.mean() * 1e3), "max_ms": float(x.max() * 1e3)}


def run_served(seed: int = 0, load: float = 0.8, target_s: float = 8.0) -> dict:
    """Serving: Poisson arrivals at load / E[plain service time]; each request 2D or 3D (equal
    odds), size log-uniform [1e3, 1e6], uniform in the unit square / cube.  The server (two
    engines, continuous batching, admission by device-memory footprint) is driven in real time
    by a submitting thread; latency = result ready - submission, on the library's clock.  The
    plain path serves the same stream FIFO one request at a time: its service times are
    measured back to back and its latencies follow from the arrival times.  A saturation run
    (every request queued at once) gives both forms' throughput.  Every request's simplex set
    is compared between the two (order-free hash)."""
    import threading
    from voronoids_amd import batch as B
    from voronoids_amd import datasets
    lo, hi = 1000, 1_000_000
    # service-time model of the plain path, per dimension: a + b * n from four sizes
    fit = {}
    for dim in (2, 3):
        ns, ts = [], []
        for n in (1000, 10_000, 100_000, 1_000_000):
            Pp = datasets.uniform(n, dim, seed=seed + 7, stream=dim)
            B.plain_simplices(Pp, 1)                           # warm
            t0 = time.perf_counter()
            B.plain_simplices(Pp, 1)
            ns.append(n)
            ts.append(time.perf_counter() - t0)
        b, a = np.polyfit(ns, ts, 1)
        fit[dim] = (float(a), float(b))
    mean_n = (hi - lo) / math.log(hi / lo)
    es = 0.5 * sum(fit[d][0] + fit[d][1] * mean_n for d in (2, 3))
    lam = load / es
    n_req = int(min(max(target_s / es, 40), 1000))
    u = datasets.unit(seed, 30, 0, 3 * n_req).reshape(3, n_req)
    sizes = np.floor(np.exp(np.log(lo) + u[0] * (np.log(hi) - np.log(lo)))).astype(np.int64)
    dims = np.where(u[1] < 0.5, 2, 3)
    arrivals = np.cumsum(-np.log1p(-u[2]) / lam)
    reqs = [datasets.uniform(int(sizes[k]), int(dims[k]), seed=seed + 31, stream=k) for k in range(n_req)]
    print(f"# serving: {n_req} requests, E[S_plain] {es * 1e3:.1f} ms, rate {lam:.1f}/s, "
          f"{int(sizes.sum())} points", file=sys.stderr, flush=True)

    def serve(arrival_times):
        srv = B.Server(max_points_2d=16_000_000, max_points_3d=16_000_000)
        try:
            for d in (2, 3):                                    # warm-up: graphs of both engines
                srv.submit(datasets.uniform(2000, d, seed=99), seed=0, tag=-d)
            got, warm = {}, 0
            while warm < 2:
                warm += sum(1 for r in srv.results(timeout_ms=500) if r.tag < 0)
            err = []

            def submitter():
                t0 = time.perf_counter()
                try:
                    for k in range(n_req):
                        w = t0 + arrival_times[k] - time.perf_counter()
                        if w > 0:
                            time.sleep(w)
                        srv.submit(reqs[k], seed=1000 + k, tag=k)
                except Exception as ex:                     # reported by the collector
                    err.append(ex)
            th = threading.Thread(target=submitter, daemon=True)
            th.start()
            hashes, lat, first, last = {}, {}, None, None
            while len(hashes) < n_req and not err:
                for r in srv.results(timeout_ms=200):
                    if r.tag < 0:
                        continue
                    hashes[r.tag] = B.simplex_set_hash(r.simplices)
                    lat[r.tag] = r.latency
                    first = r.t_submit if first is None else min(first, r.t_submit)
                    last = r.t_done if last is None else max(last, r.t_done)
            th.join()
            if err:
                raise err[0]
            stats = {d: srv.stats(d) for d in (2, 3)}
        finally:
            srv.close()
        return hashes, lat, last - first, stats

    h_srv, lat_srv, span, est = serve(arrivals)
    print(f"# served (open loop): makespan {span:.2f} s, p50 {np.percentile(list(lat_srv.values()), 50) * 1e3:.1f} ms",
          file=sys.stderr, flush=True)
    _, _, span_sat, _ = serve(np.zeros(n_req))
    # plain path: back-to-back service times, then FIFO latencies for the same arrivals
    svc = np.zeros(n_req)
    same = 0
    for k in range(n_req):
        t0 = time.perf_counter()
        s, _ = B.plain_simplices(reqs[k], 1000 + k)
        svc[k] = time.perf_counter() - t0
        same += int(B.simplex_set_hash(s) == h_srv[k])
    end = 0.0
    lat_plain = np.zeros(n_req)
    for k in range(n_req):
        start = max(arrivals[k], end)
        end = start + svc[k]
        lat_plain[k] = end - arrivals[k]
    out = {
        "workload": f"{n_req} requests, 2D/3D equal odds, sizes log-uniform [1e3, 1e6], Poisson arrivals at "
                    f"{load:.0%} of the plain path's capacity ({lam:.1f}/s)",
        "served": {**_percentiles(list(lat_srv.values())), "sets_per_s": n_req / span,
                   "saturated_sets_per_s": n_req / span_sat},
        "plain_fifo": {**_percentiles(lat_plain), "sets_per_s": n_req / max(end - arrivals[0], 1e-9),
                       "saturated_sets_per_s": n_req / svc.sum(), "mean_service_ms": float(svc.mean() * 1e3)},
        "identical_output": {"compared_requests": n_req, "identical": same},
        "engines": {str(d): {k: est[d][k] for k in ENGINE_STAT_KEYS} for d in (2, 3)},
    }
    out["saturated_speedup"] = out["served"]["saturated_sets_per_s"] / out["plain_fifo"]["saturated_sets_per_s"]
    out["p99_ratio_plain_over_served"] = out["plain_fifo"]["p99_ms"] / out["served"]["p99_ms"]
    return out


def _tier_fractions(st: dict) -> dict:
    out = {}
    for i, name in enumerate(("orient", "insphere")):
        tot = st["tier1"][i] + st["tier2"][i] + st["tier3"][i] + st["deferred"][i]
        out[name] = {"calls": tot, "fp32": st["tier1"][i] / max(tot, 1), "fp64_filter": st["tier2"][i] / max(tot, 1),
                     "fp64_expansion": (st["tier3"][i] + st["deferred"][i]) / max(tot, 1)}
    return out


def run_precision(sample: int = 20000) -> dict:
    """Precision tiers on the two hard 3D inputs: (a) n = 4,000,000 clustered (80% from 64
    Gaussians, sigma 0.02, 20% uniform), (b) the 128^3 lattice with +-1e-9 jitter.  Each is
    triangulated once (per-tier resolution fractions of every predicate call, exact device
    check of the result), and `sample` predicate instances of the same kinds of geometry go
    through the tier cascade one by one (vo_predicates) and through the oracle's exact
    big-integer predicates: the sign disagreements are counted (expected: none)."""
    import voronoids_amd as vo
    from oracle import oracle as O
    from voronoids_amd import batch as B
    from voronoids_amd import datasets
    out = {}
    for key, P in (("clustered_4m", datasets.clustered(4_000_000, seed=0)), ("lattice_128", datasets.lattice_jitter(128, seed=0))):
        t = vo.DelaunayTree(P)
        t.stage(P, seed=0)
        t.run()
        st = t.stats()
        viol = t.violations()
        t.close()
        out[key] = {"n": P.shape[0], "ms": st["ms_run"], "pts_per_s": P.shape[0] / (st["ms_run"] / 1e3),
                    "rounds": st["rounds"], "tiers": _tier_fractions(st), "exact_evaluated": st["exact_evaluated"],
                    "solo_runs": st["solo_runs"], "violations": viol}
        print(f"# precision {key}: {st['ms_run']:.1f} ms, violations {viol}", file=sys.stderr, flush=True)
    fn = {"orient3d": O.orient3d, "insphere": O.insphere}
    check = {}
    for fam in ("clustered", "lattice_jitter", "rounded_sphere", "exact_degenerate"):
        for name in ("orient3d", "insphere"):
            T = datasets.PREDICATE_FAMILIES[fam](name, sample // 8, seed=7)
            signs, tiers = B.predicates(name, T)
            ref = np.array([fn[name](*T[i]) for i in range(T.shape[0])])
            check[f"{fam}/{name}"] = {"n": int(T.shape[0]), "disagreements": int((signs != ref).sum()),
                                      "tier_counts": [int((tiers == k).sum()) for k in (1, 2, 3)]}
    out["sign_check_vs_exact_oracle"] = check
    return out


def copy_calibration(nbytes: int, reps: int = 3) -> None:
    """Device-to-device copies of a known size (the runtime's copy kernel), so that rocprofv3's
    FETCH_SIZE / WRITE_SIZE can be scaled to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="3d4m", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-sets measurement")
    ap.add_argument("--no-served", action="store_true", help="skip the served-requests measurement")
    ap.add_argument("--no-precision", action="store_true", help="skip the precision-tier measurement")
    ap.add_argument("--only", choices=["batched", "served", "precision"], help="run only that measurement")
    ap.add_argument("--check", action="store_true", help="run the exact device check after the last step")
    ap.add_argument("--kernel-timing-steps", type=int, default=-1, metavar="K",
                    help="further steps with HIP events around every round's kernels, for the roofline "
                         "(default: as many as --steps; 0: none, e.g. under a profiler)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (for rocprofv3 counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is designed but not built or measured", file=sys.stderr)
        return 2

    if args.only:
        r = {"batched": run_batched, "served": run_served, "precision": run_precision}[args.only]()
        print(json.dumps({"metric": args.only, **r}), flush=True)
        return 0

    import voronoids_amd as vo
    w = WORKLOADS[args.workload]
    P = make_points(w)
    n, dim = P.shape[0], P.shape[1]
    print(f"# workload {args.workload}: {w['label']}", file=sys.stderr, flush=True)

    # timed steps as shipped (no per-round events), then as many steps with HIP events around
    # every round's cavity and commit kernels: the per-kernel times of the roofline
    runs, runs_t = [], []
    last = None
    kts = args.steps if args.kernel_timing_steps < 0 else args.kernel_timing_steps
    for i in range(args.warmup + args.steps + kts):
        timed_kernels = i >= args.warmup + args.steps
        t = vo.DelaunayTree(P)
        if timed_kernels:
            t.config("round_timing", 1)
        t.stage(P, seed=0)
        t0 = time.perf_counter()
        t.run()                                   # synchronizes the device before returning
        wall = time.perf_counter() - t0
        st = t.stats()
        phase = "warmup" if i < args.warmup else ("kernel-timed step" if timed_kernels else "step")
        print(f"# {phase} {i}: run {st['ms_run']:.1f} ms (wall {wall * 1e3:.1f} ms), rounds {st['rounds']}, "
              f"cavity {st['ms_locate_cavity']:.1f} ms, commit {st['ms_commit']:.1f} ms",
              file=sys.stderr, flush=True)
        if i >= args.warmup:
            (runs_t if timed_kernels else runs).append(st)
        if last is not None:
            last.close()
        last = t
    viol = None
    if args.check:
        viol = last.violations()
    last.close()

    ms = float(np.mean([r["ms_run"] for r in runs]))
    value = n / (ms / 1e3)
    st = (runs_t or runs)[-1]
    alg = algorithmic_bytes(dim, n, st)
    kt = {"k_cavity": float(np.mean([r["ms_locate_cavity"] for r in runs_t])) if runs_t else 0.0,
          "k_commit": float(np.mean([r["ms_commit"] for r in runs_t])) if runs_t else 0.0}
    launches = {"k_cavity": st["launches_locate_cavity"], "k_commit": st["launches_commit"]}
    dom = max(kt, key=kt.get)
    achieved = alg[dom] / (kt[dom] / 1e3) / 1e9 if kt[dom] > 0 else None
    traffic = measured_traffic(args.workload, dom)
    tot_calls = [st["tier1"][i] + st["tier2"][i] + st["tier3"][i] + st["deferred"][i] for i in range(2)]
    tiers = {name: {"fp32": st["tier1"][i] / max(tot_calls[i], 1), "fp64_filter": st["tier2"][i] / max(tot_calls[i], 1),
                    "fp64_expansion": st["tier3"][i] / max(tot_calls[i], 1),
                    "deferred": st["deferred"][i] / max(tot_calls[i], 1), "calls": tot_calls[i]}
             for i, name in enumerate(("orient", "insphere"))}
    out = {
        "metric": "3D Delaunay inserted pts/sec @4M; batched sets/s + p50/p99 latency vs plain GPU"
        if args.workload == "3d4m" else f"{dim}D Delaunay inserted pts/sec",
        "value": value, "unit": "pts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["label"], "n": n, "dim": dim, "timed": "vo_tri_run, inputs resident in HBM"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic["bytes_per_launch"],
                     "traffic_source": traffic,
                     "alg_bytes_per_launch": alg[dom] / max(launches[dom], 1),
                     "avg_launch_ms": kt[dom] / max(launches[dom], 1), "launches": launches[dom],
                     "kernel_times": (f"HIP events around every round's kernels in {len(runs_t)} further steps "
                                      f"({float(np.mean([r['ms_run'] for r in runs_t])):.1f} ms per step with the events)")
                                     if runs_t else None},
        "kernels_ms": kt, "rounds": st["rounds"], "tiers": tiers,
        "events": {k: st[k] for k in ("points_inserted", "killed", "created", "walk_steps", "chain_hops",
                                      "max_killed", "max_facets", "windows_total", "exact_evaluated")},
    }
    if viol is not None:
        out["violations"] = viol
    if args.workload == "3d4m" and not args.no_batched:
        out["batched"] = run_batched()
    if args.workload == "3d4m" and not args.no_served:
        out["served"] = run_served()
    if args.workload == "3d4m" and not args.no_precision:
        out["precision"] = run_precision()
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(dim, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
