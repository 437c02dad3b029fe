"""Benchmark: T-DFS subgraph counting on one MI355X, batched vs one pattern per launch.

Headline workload (default "c4"): the 64 seeded patterns of
tests/golden/c4_patterns.json (3-6 vertices, half labeled with 4 labels; rule
in tdfs_amd/workloads.py) on the labeled Chung-Lu graph n = 1.1M, m = 3.0M,
gamma = 2.1, seed 2 (labels seed 3), resident in HBM, exact counts.

One step = every pattern counted once over the whole graph.
  batched : all patterns in ONE persistent launch, replayed from a hipGraph
  plain   : one launch per pattern (tdfs_count), the reference's form
value = total matches / batched step time (matches/s, higher is better).

Also reported: the plain form, the c2 workload (4-clique + diamond, unlabeled),
the roofline of the match kernel (algorithmic bytes counted by the kernel
itself / hipEvent kernel time), and the CPU oracle on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--workload c4|c2|c3|c1] [--no-cpu] [--no-c2]
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
# timed match kernel per tdfs_tuning.waves_per_simd
KERNEL_NAMES = {0: "tdfs_match_kernel", 4: "tdfs_match_kernel_w4", 5: "tdfs_match_kernel_w5",
                6: "tdfs_match_kernel_w6", 8: "tdfs_match_kernel_w8"}


def load_workload(name: str):
    """(graph, {name: (.g text, labeled)}, description)"""
    from tdfs_amd import workloads as W

    g = W.graph(name)
    if name == "c4":
        d = json.loads((ROOT / "tests" / "golden" / "c4_patterns.json").read_text())
        pats = {f"p{p['index']:02d}_k{p['k']}{'L' if p['labeled'] else ''}": (p["text"], p["labeled"])
                for p in d["patterns"]}
        desc = ("64 seeded patterns (k 3-6, 32 labeled) on Chung-Lu n=1.1M m=3.0M gamma=2.1 seed 2, "
                "4 labels seed 3")
    else:
        pats = W.patterns(name)
        desc = {"c1": "Erdos-Renyi n=20000 m=100000 seed 1; triangle",
                "c2": "Chung-Lu n=1.1M m=3.0M gamma=2.1 dmax~30k seed 2, unlabeled; 4-clique + diamond",
                "c3": "c2 graph + 4 labels seed 3; labeled house + 5-clique"}[name]
    return g, pats, desc


def cpu_baseline(g, pats, threads: int, budget_s: float = 15.0) -> dict:
    """The oracle on seeded random windows of level-0 edges, 16 windows at a
    time, until budget_s of wall time has been spent; matches/s = sampled
    matches / time spent."""
    from oracle import oracle as O

    rng = np.random.default_rng(1234)
    win = 256
    oplans = [O.plan(t, lab) for t, lab in pats.values()]
    prep = O.Prepared(g.rowptr, g.colidx, g.labels)
    order = rng.permutation(max(1, g.nnz // win))
    matches, spent, used = 0, 0.0, 0
    while spent < budget_s and used < order.shape[0]:
        starts = np.sort(order[used:used + 16]) * win
        used += starts.shape[0]
        t0 = time.perf_counter()
        for p in oplans:
            matches += prep.count_ranges(p, starts, starts + win, threads=threads)["count"]
        spent += time.perf_counter() - t0
    prep.close()
    edges = used * win
    return {
        "value": matches / spent,
        "unit": "matches/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{used} seeded random windows of {win} level-0 edges "
                  f"({100.0 * edges / g.nnz:.2f}% of the directed edges), all {len(oplans)} patterns, "
                  f"oracle/tdfs_oracle.c on {threads} threads, {spent:.1f} s",
    }


_HIP = None


def device_sync() -> None:
    """Whole-device synchronisation around the timed region (every stream of
    the library included): hipDeviceSynchronize from the HIP runtime the
    library already loaded (torch.cuda.synchronize() without importing torch)."""
    global _HIP
    if _HIP is None:
        import ctypes

        try:  # the soname libtdfs.so links, so the already-loaded runtime is reused
            _HIP = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        except OSError:
            _HIP = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
    rc = _HIP.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


def measure(dev, dg, pats, steps: int, warmup: int, plain_too: bool = True) -> dict:
    """Exact counts + algorithmic bytes (statistics launch), then timed
    batched (hipGraph replay) and plain (one launch per pattern) steps."""
    import tdfs_amd as T

    names = list(pats)
    plans = [T.plan_pattern(t, lab) for t, lab in pats.values()]
    dev.set_stats(True)
    st = dev.count_batch(dg, plans)
    prof = dev.last_profile()
    dev.set_stats(False)
    counts = [int(r.count) for r in st]
    alg_bytes = sum(int(r.alg_bytes) for r in st)
    batch = T.Batch(dev, dg, plans)

    def batched():
        res = batch.run()
        got = [int(r.count) for r in res]
        assert got == counts, "batched counts differ from the statistics launch"
        return res[0].kernel_ms

    def plain():
        ms = 0.0
        for p, c in zip(plans, counts):
            r = dev.count(dg, p)
            assert int(r.count) == c
            ms += r.kernel_ms
        return ms

    for _ in range(warmup):
        batched()
        if plain_too:
            plain()
    kb, kp = [], []
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        kb.append(batched())
    device_sync()
    t_b = (time.perf_counter() - t0) / steps
    t_p = float("nan")
    if plain_too:
        t0 = time.perf_counter()
        for _ in range(steps):
            kp.append(plain())
        device_sync()
        t_p = (time.perf_counter() - t0) / steps
    batch.close()
    total = sum(counts)
    kernel_ms = float(np.mean(kb))
    return {
        "names": names, "counts": dict(zip(names, counts)), "total": total, "alg_bytes": alg_bytes,
        "t_batched": t_b, "t_plain": t_p, "kernel_ms_batched": kernel_ms,
        "kernel_ms_plain": float(np.mean(kp)) if kp else float("nan"),
        "achieved_gbs": alg_bytes / (kernel_ms * 1e-3) / 1e9, "profile": prof,
    }


def per_pattern(dev, dg, pats) -> list[dict]:
    """One launch per pattern (tdfs_count, the tuned kernel) and one statistics
    launch per pattern: kernel ms, count and the cycle split of each pattern,
    slowest first."""
    import tdfs_amd as T

    rows = []
    for name, (text, lab) in pats.items():
        plan = T.plan_pattern(text, lab)
        dev.set_stats(False)
        r = dev.count(dg, plan)
        dev.set_stats(True)
        dev.count(dg, plan)
        prof = dev.last_profile()
        dev.set_stats(False)
        tot = max(1, prof["total"])
        rows.append({"pattern": name, "k": int(plan.k), "labeled": bool(plan.labeled), "kernel_ms": r.kernel_ms,
                     "count": int(r.count),
                     "cycles_frac": {k: round(prof[k] / tot, 3) for k in ("filter", "extend", "leaf", "idle")}})
    rows.sort(key=lambda x: -x["kernel_ms"])
    return rows


def nan_none(x: float):
    return None if x != x else x


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--per-pattern", action="store_true",
                    help="instead of the bench line: each pattern alone (ms, count, cycle split), one JSON line each")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one-launch-per-pattern form (profiled runs: the match kernel then runs "
                         "only as the batched launch)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (a kernel of known traffic "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only. Across GPUs, T-DFS would shard level-0 edge chunks "
              "(tdfs_count's [edge_begin, edge_end) range, chunks interleaved over GPUs, graph replicated) "
              "and finish with one 8-byte sum per pattern; that multi-GPU driver is not part of this library.",
              file=sys.stderr)
        return 2

    import tdfs_amd as T

    tfile = ROOT / "profiles" / "autotune.json"
    tuned_all = json.loads(tfile.read_text()) if tfile.exists() else {}
    tuned = tuned_all.get(args.workload, {}).get("best", {})

    if args.copy_calibration > 0:
        import torch

        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy calibration: {n} bytes read, {n} bytes written", file=sys.stderr)
        del src, dst

    g, pats, desc = load_workload(args.workload)
    dev = T.Device(0)
    if tuned:
        dev.set_tuning(**tuned)
    dg = dev.upload(g)
    if args.per_pattern:
        for row in per_pattern(dev, dg, pats):
            print(json.dumps(row), flush=True)
        dg.free()
        dev.close()
        return 0
    m = measure(dev, dg, pats, args.steps, args.warmup, plain_too=not args.no_plain)
    dg.free()

    traffic = None
    trfile = ROOT / "profiles" / "traffic.json"
    if trfile.exists():
        traffic = json.loads(trfile.read_text()).get(args.workload, {}).get("hbm_bytes_per_launch")

    extra = {}
    if args.workload != "c2" and not args.no_c2:
        g2, p2, d2 = load_workload("c2")
        dev.set_tuning(**{k: 0 for k in tuned})  # zero = library default
        dev.set_tuning(**tuned_all.get("c2", {}).get("best", {}))
        dg2 = dev.upload(g2)
        m2 = measure(dev, dg2, p2, args.steps, args.warmup)
        dg2.free()
        extra = {"c2": {"workload": d2, "counts": m2["counts"], "batched_matches_per_s": m2["total"] / m2["t_batched"],
                        "plain_matches_per_s": m2["total"] / m2["t_plain"], "ms_per_step": m2["t_batched"] * 1e3,
                        "kernel_ms_batched": m2["kernel_ms_batched"],
                        "achieved_alg_gbs": m2["achieved_gbs"],
                        "tuning": tuned_all.get("c2", {}).get("best", "defaults")}}

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(g, pats, threads)

    prof = m["profile"]
    tot = max(1, prof["total"])
    line = {
        "metric": "matches/sec (batched launch vs one pattern per launch)",
        "value": m["total"] / m["t_batched"],
        "unit": "matches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["t_batched"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded in-repo generator)",
        "config": {"workload": f"{args.workload}: {desc}", "patterns": len(m["names"]),
                   "total_matches": m["total"], "graph": {"n": g.n, "nnz": g.nnz, "dmax": g.dmax},
                   "tuning": tuned or "defaults"},
        "plain_matches_per_s": nan_none(m["total"] / m["t_plain"]),
        "plain_ms_per_step": nan_none(m["t_plain"] * 1e3),
        "batched_speedup": nan_none(m["t_plain"] / m["t_batched"]),
        "kernel_ms_batched": m["kernel_ms_batched"],
        "kernel_ms_plain_sum": nan_none(m["kernel_ms_plain"]),
        "roofline": {"bound": "hbm", "achieved": m["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": m["achieved_gbs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": m["alg_bytes"],
                     "kernel": KERNEL_NAMES.get(int(tuned.get("waves_per_simd", 0)), "tdfs_match_kernel")},
        "phase_cycles_frac": {k: round(v / tot, 4) for k, v in prof.items() if k not in ("total", "waves")},
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line))
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
