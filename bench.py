"""Benchmark: T-DFS subgraph counting on one MI355X.

Workload (default "c2"): unlabeled Chung-Lu graph, n = 1.1M, m = 3.0M
undirected edges, degree exponent 2.1, largest expected degree 30k, seed 2,
resident in HBM; patterns 4-clique and diamond (the reference's
data/pattern/2.g and 1.g), exact counts.

One step = every pattern counted once over the whole graph.
  batched : all patterns in ONE persistent launch (tdfs_count_batch)
  plain   : one launch per pattern (tdfs_count), the reference's form
value = total matches / batched step time  (matches/s, higher is better).

Also reported: the plain form's matches/s, the roofline of the match kernel
(algorithmic bytes counted by the kernel itself / hipEvent kernel time), and
the CPU oracle on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--workload c2|c1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

TRIANGLE = "v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 0 2 1\ne 1 2 1\n"


def workload(name: str):
    import tdfs_amd as T

    pat = ROOT / "tests" / "golden" / "patterns"
    if name == "c2":
        g = T.CSRGraph.chung_lu(1_100_000, 3_000_000, 2.1, 30000.0, 2)
        pats = {"4-clique": (pat / "2.g").read_text(), "diamond": (pat / "1.g").read_text()}
        desc = "Chung-Lu n=1.1M m=3.0M gamma=2.1 dmax~30k seed 2, unlabeled; 4-clique + diamond"
    elif name == "c1":
        g = T.CSRGraph.erdos_renyi(20_000, 100_000, 1)
        pats = {"triangle": TRIANGLE}
        desc = "Erdos-Renyi n=20000 m=100000 seed 1, unlabeled; triangle"
    else:
        raise SystemExit(f"unknown workload {name}")
    plans = {k: T.plan_pattern(v, 0) for k, v in pats.items()}
    return g, pats, plans, desc


def cpu_baseline(g, pats, threads: int, budget_s: float = 15.0) -> dict:
    """The oracle on seeded random windows of level-0 edges: the sample is
    doubled until one pass over it takes about budget_s / 4 of wall time;
    matches/s = sampled matches / time of the final pass."""
    from oracle import oracle as O

    rng = np.random.default_rng(1234)
    win = 256
    oplans = {k: O.plan(v, 0) for k, v in pats.items()}
    nwin = 16
    while True:
        starts = np.sort(rng.choice(max(1, g.nnz // win), size=min(nwin, g.nnz // win), replace=False)) * win
        t0 = time.perf_counter()
        matches = 0
        for p in oplans.values():
            matches += O.count_ranges(g.rowptr, g.colidx, None, p, starts, starts + win, threads=threads)["count"]
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or nwin >= g.nnz // win:
            break
        nwin *= 2
    edges = int(starts.shape[0]) * win
    return {
        "value": matches / dt,
        "unit": "matches/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{starts.shape[0]} seeded random windows of {win} level-0 edges "
                  f"({100.0 * edges / g.nnz:.2f}% of the directed edges), every pattern, "
                  f"oracle/tdfs_oracle.c on {threads} threads, {dt:.1f} s",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: T-DFS shards level-0 edges across GPUs with a single "
              "8-byte count reduction per pattern, which this build does not run (see DESIGN.md).",
              file=sys.stderr)
        return 2

    import tdfs_amd as T

    g, pats, plans, desc = workload(args.workload)
    dev = T.Device(0)
    dg = dev.upload(g)
    names = list(plans)
    plist = [plans[k] for k in names]

    # statistics launch: exact counts + algorithmic bytes per pattern
    dev.set_stats(True)
    stats = {k: dev.count(dg, plans[k]) for k in names}
    dev.set_stats(False)
    counts = {k: int(stats[k].count) for k in names}
    alg_bytes = sum(int(stats[k].alg_bytes) for k in names)
    total_matches = sum(counts.values())

    def batched():
        res = dev.count_batch(dg, plist)
        for k, r in zip(names, res):
            assert int(r.count) == counts[k], (k, r.count, counts[k])
        return res[0].kernel_ms

    def plain():
        ms = 0.0
        for k in names:
            r = dev.count(dg, plans[k])
            assert int(r.count) == counts[k]
            ms += r.kernel_ms
        return ms

    for _ in range(args.warmup):
        batched()
        plain()
    kb, kp = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kb.append(batched())
    t_batched = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kp.append(plain())
    t_plain = (time.perf_counter() - t0) / args.steps
    kernel_ms = float(np.mean(kb))
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9

    traffic = None
    tfile = ROOT / "profiles" / f"traffic_{args.workload}.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(g, pats, threads)

    line = {
        "metric": "matches/sec (batched launch)",
        "value": total_matches / t_batched,
        "unit": "matches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_batched * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded in-repo generator)",
        "config": {"workload": f"{args.workload}: {desc}", "patterns": names, "counts": counts,
                   "graph": {"n": g.n, "nnz": g.nnz, "dmax": g.dmax}},
        "plain_matches_per_s": total_matches / t_plain,
        "plain_ms_per_step": t_plain * 1e3,
        "batched_speedup": t_plain / t_batched,
        "kernel_ms_batched": kernel_ms,
        "kernel_ms_plain": float(np.mean(kp)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes, "kernel": "tdfs_match_kernel"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    dg.free()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
