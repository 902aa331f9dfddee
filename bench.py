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

# arithmetic operations (add, sub, mul) of one predicate evaluation in the filter tiers: the
# stage-A expression, its permanent and the error bound (voronoids_amd/csrc/predicates.hpp); the
# same count for the FP32 tier and the FP64 filter.  Secondary evidence against the vector peaks.
PRED_FLOPS = {3: {"orient": 40, "insphere": 130}, 2: {"orient": 16, "insphere": 58}}
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6

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


TRAFFIC_PROFILES = {"3d4m": "traffic_round_kernels_3d4m_uniform.md", "2d10m": "traffic_round_kernels_2d10m_uniform.md"}


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


def cpu_threads() -> int:
    """host threads for the CPU baseline: OMP_NUM_THREADS when set (the GPU host's share for
    one GPU), else every core"""
    v = os.environ.get("OMP_NUM_THREADS", "")
    return int(v) if v.isdigit() and int(v) > 0 else (os.cpu_count() or 1)


def cpu_baseline(P: np.ndarray, threads: int) -> dict:
    """The oracle's restatement of voronoids.delaunay (src/lib.rs:104-125) with the
    reference's own arithmetic: 1e5 sequential inserts, then add_points_to_tree (make_queue
    / find_placement / level batches) in parallel on `threads` OpenMP threads; the workload's
    own points, one run (tens of seconds at n = 4M)."""
    from oracle import oracle as O
    n, dim = P.shape
    t0 = time.perf_counter()
    tr = O.Tree.delaunay(P, O.MODE_REF, nthreads=threads)
    dt = time.perf_counter() - t0
    del tr
    return {"value": n / dt, "unit": "pts/s", "cores": threads, "kind": "port",
            "sample": f"the workload itself: {dim}D uniform n={n} (1e5 sequential insert_point + {n - 100000} via "
                      f"add_points_to_tree, the reference's arithmetic), one run, {dt:.1f} s wall"}


def tuning_applied() -> dict:
    """the autotuner's stored choice per dimension, applied to every run here: its build constants
    are compiled into the library and DelaunayTree applies its runtime setting"""
    path = os.path.join(ROOT, "voronoids_amd", "tuning.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        table = json.load(f)
    return {d: {k: table[d].get(k) for k in ("build", "setting", "ms", "default_ms", "workload")} for d in sorted(table)}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary of this
    workload (FETCH_SIZE scaled by what the counter reads for random 128 B-line gathers of known
    size -- the round kernels' access pattern; a streaming copy reads twice what FETCH_SIZE
    reports -- and WRITE_SIZE by what it reads for a device-to-device copy), or None when no
    summary exists."""
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
    pts_s = int(sum(sizes[k] for k in idx))

    def plain_run(fn):
        t0 = time.perf_counter()
        out = [fn(P[off[k]:off[k + 1]], seed + k)[0] for k in idx]
        return out, time.perf_counter() - t0

    # the one-set path as a caller would run it set after set: one handle reset per set (no
    # allocation per set), and, for reference, a new handle created and destroyed per set
    plain = B.PlainPath()
    try:
        plain.simplices(P[off[0]:off[1]], seed)                  # warm-up: the handle's buffers
        simp_r, dtr = plain_run(plain.simplices)
    finally:
        plain.close()
    simp_c, dtc = plain_run(B.plain_simplices)
    same = sum(B.simplex_set_hash(res[k].simplices) == B.simplex_set_hash(a) == B.simplex_set_hash(b)
               for k, a, b in zip(idx, simp_r, simp_c))
    print(f"# plain path: {len(idx)} sets, {pts_s} points: reused handle {dtr:.2f} s, create/destroy {dtc:.2f} s; "
          f"identical {same}/{len(idx)}", file=sys.stderr, flush=True)
    ok = all(r.status == 0 for r in res)
    return {
        "workload": "4,096 3D sets, sizes log-uniform [64, 65536], i.i.d. uniform [0,1)^3 f64",
        "sets": len(res), "points": int(off[-1]), "s": dt, "sets_per_s": len(res) / dt, "pts_per_s": int(off[-1]) / dt,
        "plain": {"sets_timed": len(idx), "points_timed": pts_s, "s": dtr, "sets_per_s": len(idx) / dtr,
                  "pts_per_s": pts_s / dtr, "sample": f"every {plain_stride}th set",
                  "handle": "one vo_tri handle reset per set (vo_tri_reset), simplices exported to the host"},
        "plain_create_destroy": {"s": dtc, "sets_per_s": len(idx) / dtc, "pts_per_s": pts_s / dtc,
                                 "handle": "vo_tri_create + insert + export + vo_tri_destroy per set"},
        "speedup_sets_per_s": (len(res) / dt) / (len(idx) / dtr),
        "speedup_pts_per_s": (int(off[-1]) / dt) / (pts_s / dtr),
        "speedup_vs_create_destroy": (len(res) / dt) / (len(idx) / dtc),
        "identical_output": {"compared_sets": len(idx), "identical": int(same), "all_status_ok": ok},
        "engine": {k: st[k] - (base[k] if k not in ("max_sets_resident",) else 0) for k in ENGINE_STAT_KEYS},
    }


def _percentiles(x) -> dict:
    x = np.asarray(x, dtype=np.float64)
    return {"p50_ms": float(np.percentile(x, 50) * 1e3), "p99_ms": float(np.percentile(x, 99) * 1e3),
            "mean_ms": float(x.mean() * 1e3), "max_ms": float(x.max() * 1e3)}


def run_served(seed: int = 0, load: float = 0.8, target_s: float = 8.0) -> dict:
    """Serving: Poisson arrivals; each request 2D or 3D (equal odds), size log-uniform [1e3, 1e6],
    uniform in the unit square / cube.  Two load points:
      (a) load x the plain path's capacity (its mean service time measured here): the server (two
          engines, continuous batching, admission by device-memory footprint) and the plain path
          (one request at a time, FIFO, one reused handle per dimension) each serve the stream in
          real time, driven by a submitting thread; latency = result ready - submission;
      (b) load x the server's own saturated rate (every request queued at once), served by the
          server only (the plain path cannot keep up with it).
    Every request's simplex set is compared between the two forms at (a) (order-free hash)."""
    import queue
    import threading
    from voronoids_amd import batch as B
    from voronoids_amd import datasets
    lo, hi = 1000, 1_000_000
    plain = B.PlainPath()
    # the request stream (about target_s seconds of plain service at the mean size), then the
    # plain path's service time of every request, back to back: the load point's rate is
    # load / their mean
    for dim in (2, 3):
        plain.simplices(datasets.uniform(10_000, dim, seed=seed + 7, stream=dim), 1)      # warm
    mean_n = (hi - lo) / math.log(hi / lo)
    n_req = int(min(max(target_s / (mean_n * 1.3e-7), 40), 600))
    u = datasets.unit(seed, 30, 0, 3 * n_req).reshape(3, n_req)
    sizes = np.floor(np.exp(np.log(lo) + u[0] * (np.log(hi) - np.log(lo)))).astype(np.int64)
    dims = np.where(u[1] < 0.5, 2, 3)
    reqs = [datasets.uniform(int(sizes[k]), int(dims[k]), seed=seed + 31, stream=k) for k in range(n_req)]
    svc_b2b = np.zeros(n_req)
    for k in range(n_req):
        t0 = time.perf_counter()
        plain.simplices(reqs[k], 1000 + k)
        svc_b2b[k] = time.perf_counter() - t0
    es = float(svc_b2b.mean())
    lam = load / es
    arrivals = np.cumsum(-np.log1p(-u[2]) / lam)
    print(f"# serving: {n_req} requests, E[S_plain] {es * 1e3:.1f} ms (back to back), rate {lam:.1f}/s, "
          f"{int(sizes.sum())} points", file=sys.stderr, flush=True)

    def drive(arrival_times, submit):
        """call submit(k) at each arrival time on a thread of its own"""
        err = []

        def submitter():
            t0 = time.perf_counter()
            try:
                for k in range(len(arrival_times)):
                    w = t0 + arrival_times[k] - time.perf_counter()
                    if w > 0:
                        time.sleep(w)
                    submit(k)
            except Exception as ex:                             # reported by the caller
                err.append(ex)
        th = threading.Thread(target=submitter, daemon=True)
        th.start()
        return th, err

    def serve(arrival_times):
        m = len(arrival_times)
        srv = B.Server(max_points_2d=16_000_000, max_points_3d=16_000_000)
        try:
            for d in (2, 3):                                    # warm-up: graphs of both engines
                srv.submit(datasets.uniform(2000, d, seed=99), seed=0, tag=-d)
            warm = 0
            while warm < 2:
                warm += sum(1 for r in srv.results(timeout_ms=500) if r.tag < 0)
            th, err = drive(arrival_times, lambda k: srv.submit(reqs[k % n_req], seed=1000 + k, tag=k))
            hashes, lat, first, last = {}, {}, None, None
            while len(hashes) < m and not err:
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

    def serve_plain(arrival_times):
        """the plain path as a FIFO server in real time: requests queue as they arrive, one
        worker serves them one at a time; the output check (hashing each result) runs on a
        thread of its own, as the server's does, so that it does not delay the next request"""
        q = queue.Queue()
        th, err = drive(arrival_times, lambda k: q.put((k, time.perf_counter())))
        hashes, lat, svc = {}, {}, {}
        first, last = None, None
        hq = queue.Queue()

        def hasher():
            while True:
                item = hq.get()
                if item is None:
                    return
                hashes[item[0]] = B.simplex_set_hash(item[1])
        hth = threading.Thread(target=hasher, daemon=True)
        hth.start()
        for _ in range(len(arrival_times)):
            k, ta = q.get()
            t0 = time.perf_counter()
            s, _ = plain.simplices(reqs[k], 1000 + k)
            td = time.perf_counter()
            svc[k] = td - t0
            lat[k] = td - ta
            hq.put((k, s))
            first = ta if first is None else min(first, ta)
            last = td
        hq.put(None)
        hth.join()
        th.join()
        if err:
            raise err[0]
        return hashes, lat, last - first, svc

    try:
        h_srv, lat_srv, span, est = serve(arrivals)
        print(f"# served at {load:.0%} of the plain capacity: makespan {span:.2f} s, "
              f"p50 {np.percentile(list(lat_srv.values()), 50) * 1e3:.1f} ms", file=sys.stderr, flush=True)
        _, _, span_sat, _ = serve(np.zeros(n_req))
        mu_srv = n_req / span_sat
        h_pl, lat_pl, span_pl, svc = serve_plain(arrivals)
        print(f"# plain FIFO at {load:.0%}: makespan {span_pl:.2f} s, "
              f"p50 {np.percentile(list(lat_pl.values()), 50) * 1e3:.1f} ms", file=sys.stderr, flush=True)
        # the server at load x its own saturated rate: the same requests, cycled, new arrivals
        lam_b = load * mu_srv
        n_b = int(min(max(target_s * lam_b / 4, 200), 1200))
        ub = datasets.unit(seed, 32, 0, n_b)
        arr_b = np.cumsum(-np.log1p(-ub) / lam_b)
        _, lat_b, span_b, est_b = serve(arr_b)
        print(f"# served at {load:.0%} of its own capacity ({lam_b:.0f}/s): {n_b} requests, makespan {span_b:.2f} s, "
              f"p99 {np.percentile(list(lat_b.values()), 99) * 1e3:.1f} ms", file=sys.stderr, flush=True)
    finally:
        plain.close()
    same = sum(int(h_pl[k] == h_srv[k]) for k in range(n_req))
    svc_all = np.array([svc[k] for k in range(n_req)])
    out = {
        "workload": f"{n_req} requests, 2D/3D equal odds, sizes log-uniform [1e3, 1e6], Poisson arrivals at "
                    f"{load:.0%} of the plain path's capacity ({lam:.1f}/s)",
        "served": {**_percentiles(list(lat_srv.values())), "sets_per_s": n_req / span,
                   "saturated_sets_per_s": mu_srv},
        "plain_fifo": {**_percentiles(list(lat_pl.values())), "sets_per_s": n_req / span_pl,
                       "saturated_sets_per_s": n_req / svc_b2b.sum(), "mean_service_ms": float(svc_b2b.mean() * 1e3),
                       "mean_service_ms_in_stream": float(svc_all.mean() * 1e3),
                       "how": "observed: a FIFO worker serving the same arrivals in real time, one reused "
                              "vo_tri handle per dimension"},
        "served_at_own_load": {"rate_per_s": lam_b, "requests": n_b, **_percentiles(list(lat_b.values())),
                               "sets_per_s": n_b / span_b,
                               "workload": f"{n_b} requests (the same ones, cycled), Poisson arrivals at {load:.0%} of "
                                           f"the server's saturated rate"},
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


def measure_workload(wname: str, steps: int, warmup: int, kts: int, e2e_steps: int = 3, check: bool = False) -> dict:
    """Times `steps` constructions of one workload (after `warmup` untimed ones): each step
    creates and stages a tree (inputs resident in HBM), then vo_tri_run is timed with hipEvents on
    the library's stream, the device synchronized on both sides.  Then `kts` more steps with HIP
    events around every round's kernels (per-kernel times for the roofline), and `e2e_steps`
    end-to-end ones (host points to the simplices on the host: create, stage, run, export)."""
    import voronoids_amd as vo
    w = WORKLOADS[wname]
    P = make_points(w)
    n, dim = P.shape[0], P.shape[1]
    print(f"# workload {wname}: {w['label']}", file=sys.stderr, flush=True)
    runs, runs_t = [], []
    last = None
    for i in range(warmup + steps + kts):
        timed_kernels = i >= warmup + steps
        t = vo.DelaunayTree(P)
        if timed_kernels:
            t.config("round_timing", 1)
        t.stage(P, seed=0)
        t0 = time.perf_counter()
        t.run()                                   # synchronizes the device before returning
        wall = time.perf_counter() - t0
        st = t.stats()
        phase = "warmup" if i < warmup else ("kernel-timed step" if timed_kernels else "step")
        print(f"# {phase} {i}: run {st['ms_run']:.1f} ms (wall {wall * 1e3:.1f} ms), rounds {st['rounds']}, "
              f"cavity {st['ms_locate_cavity']:.1f} ms, commit {st['ms_commit']:.1f} ms",
              file=sys.stderr, flush=True)
        if i >= warmup:
            (runs_t if timed_kernels else runs).append(st)
        if last is not None:
            last.close()
        last = t
    viol = last.violations() if check else None
    last.close()
    # end to end: host points in, simplices out on the host
    e2e = []
    for i in range(e2e_steps):
        t0 = time.perf_counter()
        t = vo.DelaunayTree(P)
        t.stage(P, seed=0)
        t1 = time.perf_counter()
        t.run()
        t2 = time.perf_counter()
        s, _ = t.export()
        t3 = time.perf_counter()
        e2e.append((t3 - t0, t1 - t0, t2 - t1, t3 - t2, s.shape[0]))
        t.close()
    ms_all = [r["ms_run"] for r in runs]
    ms = float(np.median(ms_all))
    st = (runs_t or runs)[-1]
    alg = algorithmic_bytes(dim, n, st)
    kt = {"k_cavity": float(np.mean([r["ms_locate_cavity"] for r in runs_t])) if runs_t else 0.0,
          "k_commit": float(np.mean([r["ms_commit"] for r in runs_t])) if runs_t else 0.0}
    launches = {"k_cavity": st["launches_locate_cavity"], "k_commit": st["launches_commit"]}
    dom = max(kt, key=kt.get)
    achieved = alg[dom] / (kt[dom] / 1e3) / 1e9 if kt[dom] > 0 else None
    traffic = measured_traffic(wname, dom)
    tot_calls = [st["tier1"][i] + st["tier2"][i] + st["tier3"][i] + st["deferred"][i] for i in range(2)]
    tiers = {name: {"fp32": st["tier1"][i] / max(tot_calls[i], 1), "fp64_filter": st["tier2"][i] / max(tot_calls[i], 1),
                    "fp64_expansion": st["tier3"][i] / max(tot_calls[i], 1),
                    "deferred": st["deferred"][i] / max(tot_calls[i], 1), "calls": tot_calls[i]}
             for i, name in enumerate(("orient", "insphere"))}
    # predicate arithmetic of the filter tiers against the vector peaks, over the round kernels' time
    fl = PRED_FLOPS[dim]
    f32 = st["tier1"][0] * fl["orient"] + st["tier1"][1] * fl["insphere"]
    f64 = st["tier2"][0] * fl["orient"] + st["tier2"][1] * fl["insphere"]
    t_rk = (kt["k_cavity"] + kt["k_commit"]) / 1e3
    pred = None
    if t_rk > 0:
        pred = {"fp32_tflops": f32 / t_rk / 1e12, "fp32_peak_tflops": FP32_PEAK_TFLOPS,
                "fp32_frac": f32 / t_rk / 1e12 / FP32_PEAK_TFLOPS,
                "fp64_tflops": f64 / t_rk / 1e12, "fp64_peak_tflops": FP64_PEAK_TFLOPS,
                "fp64_frac": f64 / t_rk / 1e12 / FP64_PEAK_TFLOPS,
                "calls_per_s": sum(tot_calls) / t_rk,
                "flops_per_call": fl, "time": "cavity + commit kernel time of the kernel-timed steps"}
    e2e_ms = float(np.median([e[0] for e in e2e]) * 1e3) if e2e else None
    out = {
        "value": n / (ms / 1e3), "ms": ms, "ms_mean": float(np.mean(ms_all)), "ms_min": float(np.min(ms_all)),
        "n": n, "dim": dim, "label": w["label"],
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic["bytes_per_launch"],
                     "traffic_source": traffic,
                     "alg_bytes_per_launch": alg[dom] / max(launches[dom], 1),
                     "avg_launch_ms": kt[dom] / max(launches[dom], 1), "launches": launches[dom],
                     "kernel_times": (f"HIP events around every round's kernels in {len(runs_t)} further steps "
                                      f"({float(np.mean([r['ms_run'] for r in runs_t])):.1f} ms per step with the events)")
                                     if runs_t else None},
        "kernels_ms": kt, "rounds": st["rounds"], "tiers": tiers, "predicate_throughput": pred,
        "events": {k: st[k] for k in ("points_inserted", "killed", "created", "walk_steps", "chain_hops",
                                      "max_killed", "max_facets", "windows_total", "exact_evaluated", "records")},
        "end_to_end": None if not e2e else {
            "ms": e2e_ms, "pts_per_s": n / (e2e_ms / 1e3),
            "parts_ms": {"create_stage": float(np.median([e[1] for e in e2e]) * 1e3),
                         "run": float(np.median([e[2] for e in e2e]) * 1e3),
                         "export": float(np.median([e[3] for e in e2e]) * 1e3)},
            "steps": len(e2e),
            "timed": "host points to every live simplex on the host: vo_tri_create (bounding sphere), vo_tri_stage "
                     "(priority order, H2D), vo_tri_run, vo_tri_export (D2H); wall clock, median"},
    }
    if viol is not None:
        out["violations"] = viol
    return out


def gather_calibration(n_lines: int, reads: int, reps: int = 3) -> None:
    """Random gathers of a known footprint (torch's index_select on the device), so that
    rocprofv3's FETCH_SIZE can be scaled for the round kernels' access pattern: `reads` random
    128 B lines read whole (rows of 32 floats), then `reads` random 16 B reads, one per random
    line (rows of 4 floats at a line's start), out of a table of n_lines lines."""
    import torch
    dev = torch.device("cuda:0")
    x = torch.ones(n_lines * 32, dtype=torch.float32, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    idx = torch.randint(0, n_lines, (reads,), device=dev, generator=g)
    lines, quads = x.view(n_lines, 32), x.view(n_lines * 8, 4)
    for _ in range(reps):
        a = torch.index_select(lines, 0, idx)
        b = torch.index_select(quads, 0, idx * 8)
    torch.cuda.synchronize()
    print(json.dumps({"gather_calibration": {"table_bytes": n_lines * 128, "reads": reads, "reps": reps,
                                             "line_bytes_read": reads * 128, "quad_bytes_read": reads * 16,
                                             "checksum": float(a.sum() + b.sum())}}), flush=True)


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
    ap.add_argument("--no-second", action="store_true", help="skip the 2D n = 10M line of the default run")
    ap.add_argument("--only", choices=["batched", "served", "precision"], help="run only that measurement")
    ap.add_argument("--check", action="store_true", help="run the exact device check after the last step")
    ap.add_argument("--kernel-timing-steps", type=int, default=-1, metavar="K",
                    help="further steps with HIP events around every round's kernels, for the roofline "
                         "(default: as many as --steps; 0: none, e.g. under a profiler)")
    ap.add_argument("--e2e-steps", type=int, default=3, help="end-to-end steps (host points to host simplices)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (for rocprofv3 counter scaling)")
    ap.add_argument("--gather-calibration", type=int, default=0, metavar="READS",
                    help="only run READS random 128 B-line and 16 B gathers over a 16 GiB table "
                         "(for rocprofv3 counter scaling of random access)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gather_calibration:
        gather_calibration(1 << 27, args.gather_calibration)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is designed but not built or measured", file=sys.stderr)
        return 2

    if args.only:
        r = {"batched": run_batched, "served": run_served, "precision": run_precision}[args.only]()
        print(json.dumps({"metric": args.only, **r}), flush=True)
        return 0

    kts = args.steps if args.kernel_timing_steps < 0 else args.kernel_timing_steps
    m = measure_workload(args.workload, args.steps, args.warmup, kts, args.e2e_steps, args.check)
    n, dim = m["n"], m["dim"]
    out = {
        "metric": "3D Delaunay inserted pts/sec @4M; batched sets/s + p50/p99 latency vs plain GPU"
        if args.workload == "3d4m" else f"{dim}D Delaunay inserted pts/sec",
        "value": m["value"], "unit": "pts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": m["label"], "n": n, "dim": dim,
                   "timed": "vo_tri_run, inputs resident in HBM; value = n / median step time"},
        "roofline": m["roofline"],
        "ms_mean": m["ms_mean"], "ms_min": m["ms_min"],
        "kernels_ms": m["kernels_ms"], "rounds": m["rounds"], "tiers": m["tiers"],
        "predicate_throughput": m["predicate_throughput"], "events": m["events"], "end_to_end": m["end_to_end"],
    }
    if "violations" in m:
        out["violations"] = m["violations"]
    out["tuning"] = tuning_applied()
    if args.workload == "3d4m" and not args.no_second:
        s2 = measure_workload("2d10m", args.steps, args.warmup, kts, args.e2e_steps, args.check)
        out["second_config"] = {"metric": "2D Delaunay inserted pts/sec @10M", "value": s2["value"], "unit": "pts/s",
                                "ms_per_step": s2["ms"], "config": {"workload": s2["label"], "n": s2["n"], "dim": 2},
                                **{k: s2[k] for k in ("roofline", "kernels_ms", "rounds", "tiers",
                                                      "predicate_throughput", "events", "end_to_end")}}
    if args.workload == "3d4m" and not args.no_batched:
        out["batched"] = run_batched()
    if args.workload == "3d4m" and not args.no_served:
        out["served"] = run_served()
    if args.workload == "3d4m" and not args.no_precision:
        out["precision"] = run_precision()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(make_points(WORKLOADS[args.workload]), cpu_threads())
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
