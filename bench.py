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
    ap.add_argument("--check", action="store_true", help="run the exact device check after the last step")
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

    import voronoids_amd as vo
    w = WORKLOADS[args.workload]
    P = make_points(w)
    n, dim = P.shape[0], P.shape[1]
    print(f"# workload {args.workload}: {w['label']}", file=sys.stderr, flush=True)

    runs = []
    last = None
    for i in range(args.warmup + args.steps):
        t = vo.DelaunayTree(P)
        t.stage(P, seed=0)
        t0 = time.perf_counter()
        t.run()                                   # synchronizes the device before returning
        wall = time.perf_counter() - t0
        st = t.stats()
        phase = "warmup" if i < args.warmup else "step"
        print(f"# {phase} {i}: run {st['ms_run']:.1f} ms (wall {wall * 1e3:.1f} ms), rounds {st['rounds']}, "
              f"cavity {st['ms_locate_cavity']:.1f} ms, commit {st['ms_commit']:.1f} ms",
              file=sys.stderr, flush=True)
        if i >= args.warmup:
            runs.append(st)
        if last is not None:
            last.close()
        last = t
    viol = None
    if args.check:
        viol = last.violations()
    last.close()

    ms = float(np.mean([r["ms_run"] for r in runs]))
    value = n / (ms / 1e3)
    st = runs[-1]
    alg = algorithmic_bytes(dim, n, st)
    kt = {"k_cavity": float(np.mean([r["ms_locate_cavity"] for r in runs])),
          "k_commit": float(np.mean([r["ms_commit"] for r in runs]))}
    launches = {"k_cavity": st["launches_locate_cavity"], "k_commit": st["launches_commit"]}
    dom = max(kt, key=kt.get)
    achieved = alg[dom] / (kt[dom] / 1e3) / 1e9
    traffic = measured_traffic(args.workload, dom)
    tot_calls = [st["tier1"][i] + st["tier2"][i] + st["tier3"][i] + st["deferred"][i] for i in range(2)]
    tiers = {name: {"fp32": st["tier1"][i] / max(tot_calls[i], 1), "fp64_filter": st["tier2"][i] / max(tot_calls[i], 1),
                    "fp64_expansion": st["tier3"][i] / max(tot_calls[i], 1),
                    "deferred": st["deferred"][i] / max(tot_calls[i], 1), "calls": tot_calls[i]}
             for i, name in enumerate(("orient", "insphere"))}
    out = {
        "metric": "3D Delaunay inserted pts/sec @4M" if args.workload == "3d4m" else f"{dim}D Delaunay inserted pts/sec",
        "value": value, "unit": "pts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["label"], "n": n, "dim": dim, "timed": "vo_tri_run, inputs resident in HBM"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic["bytes_per_launch"],
                     "traffic_source": traffic,
                     "alg_bytes_per_launch": alg[dom] / max(launches[dom], 1),
                     "avg_launch_ms": kt[dom] / max(launches[dom], 1), "launches": launches[dom]},
        "kernels_ms": kt, "rounds": st["rounds"], "tiers": tiers,
        "events": {k: st[k] for k in ("points_inserted", "killed", "created", "walk_steps", "chain_hops",
                                      "max_killed", "max_facets", "windows_total", "exact_evaluated")},
    }
    if viol is not None:
        out["violations"] = viol
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(dim, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
