"""Benchmark: T-DFS subgraph counting on one MI355X, batched vs one pattern per launch,
and served queries (p50 / p99 latency).

Headline workload (default "c4"): the 64 seeded patterns of
tests/golden/c4_patterns.json (k 3-7, half labeled with 4 labels, kept by a
measured feasibility rule: tests/golden/make_pattern_sets.py) on the labeled
Chung-Lu graph n = 1.1M, m = 3.0M, gamma = 2.1, seed 2 (labels seed 3),
resident in HBM, exact counts.

One step = every pattern counted once over the whole graph.
  batched : all patterns in ONE persistent launch, replayed from a hipGraph
  plain   : one launch per pattern (tdfs_count), the reference's form
value = total matches / batched step time (matches/s, higher is better).

Also reported: the plain form (timed over --plain-steps), the c2 and c3
workloads, the c5 serving leg (Poisson streams at 50-2,000 queries/s over the
n = 4M, m = 34M graph: p50 / p99 latency and matches/s), the roofline of the
match kernel (measured HBM bytes per launch from rocprofv3 counters over the
live hipEvent kernel time) and the CPU oracle on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--workload c4|c2|c3|c1] [--no-cpu] [--no-legs] [--no-serve]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from collections import Counter
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# timed match kernel per tdfs_tuning.waves_per_simd
KERNEL_NAMES = {0: "tdfs_match_kernel", 4: "tdfs_match_kernel_w4", 5: "tdfs_match_kernel_w5",
                6: "tdfs_match_kernel_w6", 8: "tdfs_match_kernel_w8"}
SERVE_RATES = (50.0, 200.0, 500.0, 1000.0, 2000.0)
SERVE_POINT = {"slots": 8, "blocks_per_cu": 1, "max_batch": 4, "waves_per_simd": 5}


def load_workload(name: str):
    """(graph, {name: (.g text, labeled)}, description, extra config)"""
    from tdfs_amd import workloads as W

    g = W.graph(name)
    extra = {}
    if name == "c4":
        d = json.loads((ROOT / "tests" / "golden" / "c4_patterns.json").read_text())
        pats = {f"p{p['index']:02d}_k{p['k']}{'L' if p['labeled'] else ''}": (p["text"], p["labeled"])
                for p in d["patterns"]}
        ks = Counter(p["k"] for p in d["patterns"])
        dens = [round((p["m"] - (p["k"] - 1)) / max(1, p["k"] * (p["k"] - 1) // 2 - (p["k"] - 1)), 2)
                if p["k"] > 2 else 1.0 for p in d["patterns"]]
        extra = {"k_histogram": {str(k): ks[k] for k in sorted(ks)},
                 "density_histogram": dict(zip(["[0,.25)", "[.25,.5)", "[.5,.75)", "[.75,1]"],
                                               np.histogram(np.minimum(dens, 0.999), bins=[0, .25, .5, .75, 1.0])[0]
                                               .tolist())),
                 "labeled": sum(p["labeled"] for p in d["patterns"]),
                 "rule": d["rule"]}
        desc = (f"{len(pats)} seeded patterns (k 3-7, half labeled) on Chung-Lu n=1.1M m=3.0M gamma=2.1 seed 2, "
                "4 labels seed 3")
    else:
        pats = W.patterns(name)
        desc = {"c1": "Erdos-Renyi n=20000 m=100000 seed 1; triangle",
                "c2": "Chung-Lu n=1.1M m=3.0M gamma=2.1 seed 2, unlabeled; 4-clique + diamond",
                "c3": "c2 graph + 4 labels seed 3; labeled house + 5-clique"}[name]
    return g, pats, desc, extra


def cpu_info() -> dict:
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def cpu_baseline(g, perm, pats, threads: int, budget_s: float = 15.0, budget_1t: float = 8.0) -> dict:
    """The oracle on seeded random windows of level-0 edges of the
    degree-ordered copy the GPU counts on, 4 windows per thread at a time, until
    budget_s of wall time has been spent (threads workers), then the same on
    one thread for budget_1t; matches/s = sampled matches / time spent."""
    from oracle import oracle as O

    H = g.relabeled(perm)
    win = 256
    oplans = [O.plan(t, lab) for t, lab in pats.values()]
    prep = O.Prepared(H.rowptr, H.colidx, H.labels)

    def run(nthreads, budget):
        rng = np.random.default_rng(1234)
        order = rng.permutation(max(1, H.nnz // win))
        matches, spent, used = 0, 0.0, 0
        while spent < budget and used < order.shape[0]:
            starts = np.sort(order[used:used + 4 * nthreads]) * win
            used += starts.shape[0]
            t0 = time.perf_counter()
            for p in oplans:
                matches += prep.count_ranges(p, starts, starts + win, threads=nthreads)["count"]
            spent += time.perf_counter() - t0
        return matches, spent, used

    m, spent, used = run(threads, budget_s)
    m1, spent1, used1 = run(1, budget_1t)
    prep.close()
    info = cpu_info()
    return {
        "value": m / spent,
        "unit": "matches/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{used} seeded random windows of {win} level-0 edges of the degree-ordered copy "
                  f"({100.0 * used * win / H.nnz:.2f}% of the directed edges), all {len(oplans)} patterns, "
                  f"oracle/tdfs_oracle.c on {threads} threads, {spent:.1f} s",
        "one_thread": {"value": m1 / spent1, "windows": used1, "seconds": round(spent1, 1)},
        "cpu_model": info["model"], "nproc": info["nproc"], "usable_cpus": info["usable"],
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


def log(*a) -> None:
    """Progress on stderr (one line per phase or step)."""
    print(f"[bench {time.strftime('%H:%M:%S')}]", *a, file=sys.stderr, flush=True)


def measure(dev, dg, pats, steps: int, warmup: int, plain_steps: int) -> dict:
    """Exact counts + algorithmic bytes (statistics launch), then timed
    batched (hipGraph replay) steps and plain (one launch per pattern) steps."""
    import tdfs_amd as T

    names = list(pats)
    plans = [T.plan_pattern(t, lab) for t, lab in pats.values()]
    dev.set_stats(True)
    t0 = time.perf_counter()
    st = dev.count_batch(dg, plans)
    log(f"statistics launch {time.perf_counter() - t0:.2f} s")
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

    for i in range(warmup):
        log(f"warm-up step {i}: {batched():.1f} ms")
    kb, kp = [], []
    device_sync()
    t0 = time.perf_counter()
    for i in range(steps):
        kb.append(batched())
        if i % 5 == 4 or i == steps - 1:
            log(f"timed step {i}: {kb[-1]:.1f} ms")
    device_sync()
    t_b = (time.perf_counter() - t0) / steps
    t_p = float("nan")
    if plain_steps > 0:
        log(f"plain warm-up step: {plain():.1f} ms")
        device_sync()
        t0 = time.perf_counter()
        for i in range(plain_steps):
            kp.append(plain())
            log(f"plain step {i}: {kp[-1]:.1f} ms")
        device_sync()
        t_p = (time.perf_counter() - t0) / plain_steps
    batch.close()
    total = sum(counts)
    kernel_ms = float(np.mean(kb))
    return {
        "names": names, "counts": dict(zip(names, counts)), "total": total, "alg_bytes": alg_bytes,
        "t_batched": t_b, "t_plain": t_p, "kernel_ms_batched": kernel_ms,
        "kernel_ms_plain": float(np.mean(kp)) if kp else float("nan"), "profile": prof,
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
        log(name, f"{r.kernel_ms:.1f} ms")
        dev.set_stats(True)
        dev.count(dg, plan)
        prof = dev.last_profile()
        ctr = dev.last_counters()
        dev.set_stats(False)
        tot = max(1, prof["total"])
        rows.append({"pattern": name, "k": int(plan.k), "labeled": bool(plan.labeled), "kernel_ms": r.kernel_ms,
                     "count": int(r.count),
                     "cycles_frac": {k: round(prof[k] / tot, 3) for k in ("filter", "extend", "leaf", "idle")},
                     "cycles": prof, "counters": ctr})
    rows.sort(key=lambda x: -x["kernel_ms"])
    return rows


def serve_leg(queries: int, rates, time_limit_ms: int) -> dict:
    """c5: the committed query stream over the resident n=4M graph, served at
    each arrival rate; p50 / p99 latency (submit -> result) and matches/s of
    the exact (complete) queries; served counts checked against the counts of
    the same queries alone."""
    import tdfs_amd as T
    from tdfs_amd import serve, workloads as W

    d = json.loads((ROOT / "tests" / "golden" / "c5_queries.json").read_text())
    qs = d["queries"][:queries]
    log("c5 serving leg")
    g = W.graph("c5")
    dev = T.Device(0)
    dg = dev.upload(g)
    plans = [T.plan_pattern(q["text"], q["labeled"]) for q in qs]
    alone = [int(dev.count(dg, p).count) for p in plans]
    # serving point: 8 slots of one workgroup per CU, the 5-wave kernel, at
    # most 4 queries per launch (swept over slots, occupancy and batch size
    # with python -m tdfs_amd.serve on this stream: 37-39 queries/s)
    dev.set_tuning(time_limit_ms=time_limit_ms, waves_per_simd=SERVE_POINT["waves_per_simd"])
    sv = serve.Server(dev, dg, slots=SERVE_POINT["slots"], blocks_per_cu=SERVE_POINT["blocks_per_cu"],
                      max_batch=SERVE_POINT["max_batch"])
    rows = []
    for rate in rates:
        r = serve.run_stream(sv, plans, rate, 50 + int(rate))
        exact = [i for i in range(len(plans)) if i not in r["incomplete"] and i not in r["failed"]]
        rows.append({"rate_qps": rate, "p50_ms": round(r["p50_ms"], 2), "p99_ms": round(r["p99_ms"], 2),
                     "throughput_qps": round(r["throughput_qps"], 1), "matches_per_s": r["matches_per_s"],
                     "incomplete": len(r["incomplete"]), "failed": len(r["failed"]),
                     "counts_equal_alone": all(r["counts"][i] == alone[i] for i in exact)})
        print(f"serve {rate:g} q/s: p50 {r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms", file=sys.stderr, flush=True)
    sv.close()
    dg.free()
    dev.close()
    ks = Counter(q["k"] for q in qs)
    return {"workload": "c5: Chung-Lu n=4M m=34M gamma=2.3 seed 5, 8 labels; "
                        f"{len(qs)} committed queries (k {dict(sorted(ks.items()))}, "
                        f"{sum(q['labeled'] for q in qs)} labeled), Poisson arrivals",
            "server": {**SERVE_POINT, "admission": "DFS-stack memory arena", "time_limit_ms": time_limit_ms},
            "streams": rows}


def reach_leg(limit_ms: int) -> list[dict]:
    """The patterns the configurations name that the benchmark does not time
    (c2's 5-cycle, c3's labeled 6-path and 6-cycle), each counted alone over
    the whole graph with a time limit: the kernel time and exact count when it
    finishes, else the limit as a lower bound on the time and the matches
    counted until then as a lower bound on the count."""
    import tdfs_amd as T
    from tdfs_amd import workloads as W

    tuned_all = json.loads((ROOT / "profiles" / "autotune.json").read_text())
    rows = []
    for wl in ("c2", "c3"):
        g = W.graph(wl)
        full = W.patterns(wl, full=True)
        dev = T.Device(0)
        dev.set_tuning(**tuned_all.get(wl, {}).get("best", {}))
        dev.set_tuning(time_limit_ms=limit_ms)
        dg = dev.upload(g)
        for name in W.OUT_OF_REACH[wl]:
            text, lab = full[name]
            log(f"{wl} {name}: counting with a {limit_ms} ms limit")
            t0 = time.perf_counter()
            r = dev.count(dg, T.plan_pattern(text, lab))
            wall = time.perf_counter() - t0
            done = not (r.flags & 1)
            rows.append({"workload": wl, "pattern": name, "labeled": bool(lab), "finished": done,
                         "kernel_ms": r.kernel_ms, "wall_s": round(wall, 2),
                         "count" if done else "count_lower_bound": int(r.count),
                         "time_limit_ms": limit_ms})
            log(json.dumps(rows[-1]))
        dg.free()
        dev.close()
    return rows


def nan_none(x: float):
    return None if x != x else x


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="steps of the one-launch-per-pattern form (after one warm-up step)")
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="skip the c2 / c3 workloads")
    ap.add_argument("--no-serve", action="store_true", help="skip the c5 serving leg")
    ap.add_argument("--serve-queries", type=int, default=200)
    ap.add_argument("--per-pattern", action="store_true",
                    help="instead of the bench line: each pattern alone (ms, count, cycle split), one JSON line each")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one-launch-per-pattern form (profiled runs: the match kernel then runs "
                         "only as the batched launch)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (a kernel of known traffic "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--tune", default="", metavar="K=V,...",
                    help="tuning fields set over the workload's tuned point (tdfs_tuning names), e.g. leaf_hash=3")
    ap.add_argument("--reach", type=int, default=0, metavar="MS",
                    help="instead of the bench line: count the patterns the benchmark does not time (c2 5-cycle, "
                         "c3 6-path and 6-cycle) alone with a time limit of MS ms, one JSON line each")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only. Across GPUs, T-DFS would shard level-0 edge chunks "
              "(tdfs_count's [edge_begin, edge_end) range, chunks interleaved over GPUs, graph replicated) "
              "and finish with one 8-byte sum per pattern; that multi-GPU driver is not part of this library.",
              file=sys.stderr)
        return 2

    import tdfs_amd as T
    from tdfs_amd import workloads as W

    if args.reach > 0:
        for row in reach_leg(args.reach):
            print(json.dumps(row), flush=True)
        return 0
    tfile = ROOT / "profiles" / "autotune.json"
    tuned_all = json.loads(tfile.read_text()) if tfile.exists() else {}
    tuned = dict(tuned_all.get(args.workload, {}).get("best", {}))
    for kv in filter(None, args.tune.split(",")):
        k, v = kv.split("=")
        tuned[k.strip()] = int(v)

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

    g, pats, desc, extra_cfg = load_workload(args.workload)
    log(f"{args.workload}: graph n={g.n} nnz={g.nnz}, {len(pats)} patterns")
    dev = T.Device(0)
    if tuned:
        dev.set_tuning(**tuned)
    dg = dev.upload(g)
    perm = dg.order()
    if args.per_pattern:
        for row in per_pattern(dev, dg, pats):
            print(json.dumps(row), flush=True)
        dg.free()
        dev.close()
        return 0
    m = measure(dev, dg, pats, args.steps, args.warmup, 0 if args.no_plain else args.plain_steps)
    dg.free()
    print(f"{args.workload}: batched {m['t_batched'] * 1e3:.1f} ms per step", file=sys.stderr, flush=True)
    kernel = KERNEL_NAMES.get(int(tuned.get("waves_per_simd", 0)), "tdfs_match_kernel")

    traffic, counters = None, None
    trfile = ROOT / "profiles" / "traffic.json"
    if trfile.exists():
        rec = json.loads(trfile.read_text()).get(args.workload, {})
        if rec.get("kernel") == kernel and rec.get("patterns") == len(m["names"]):
            traffic = rec.get("hbm_bytes_per_launch")
            counters = {k: rec[k] for k in ("l2_hit", "waves_resident", "valu_lane_util") if k in rec}
            counters["source"] = "rocprofv3 counter passes of this kernel on this workload (profiles/traffic.json)"

    legs = {}
    if not args.no_legs:
        for leg in ("c2", "c3"):
            if leg == args.workload:
                continue
            g2, p2, d2, _ = load_workload(leg)
            dev.set_tuning(**{k: 0 for k in tuned})  # zero = library default
            dev.set_tuning(**tuned_all.get(leg, {}).get("best", {}))
            dg2 = dev.upload(g2)
            m2 = measure(dev, dg2, p2, args.steps, args.warmup, args.plain_steps)
            dg2.free()
            legs[leg] = {"workload": d2, "counts": m2["counts"],
                         "batched_matches_per_s": m2["total"] / m2["t_batched"],
                         "plain_matches_per_s": nan_none(m2["total"] / m2["t_plain"]),
                         "ms_per_step": m2["t_batched"] * 1e3, "kernel_ms_batched": m2["kernel_ms_batched"],
                         "tuning": tuned_all.get(leg, {}).get("best", "defaults")}
            print(f"{leg}: batched {m2['t_batched'] * 1e3:.1f} ms per step", file=sys.stderr, flush=True)
    dev.close()

    serving = None
    if not args.no_serve and args.workload == "c4":
        serving = serve_leg(args.serve_queries, SERVE_RATES, 2000)

    cpu = None
    if not args.no_cpu:
        log("CPU baseline")
        threads = args.cpu_threads or min(16, cpu_info()["usable"])
        cpu = cpu_baseline(g, perm, pats, threads)

    prof = m["profile"]
    tot = max(1, prof["total"])
    kms = m["kernel_ms_batched"]
    achieved = traffic / (kms * 1e-3) / 1e9 if traffic else None
    line = {
        "metric": "matches/sec batched vs one-pattern-per-launch; p50/p99 query latency",
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
                   "tuning": tuned or "defaults", **extra_cfg},
        "plain_matches_per_s": nan_none(m["total"] / m["t_plain"]),
        "plain_ms_per_step": nan_none(m["t_plain"] * 1e3),
        "plain_steps": 0 if args.no_plain else args.plain_steps,
        "batched_speedup": nan_none(m["t_plain"] / m["t_batched"]),
        "kernel_ms_batched": kms,
        "kernel_ms_plain_sum": nan_none(m["kernel_ms_plain"]),
        "counts": m["counts"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "kernel": kernel,
                     "merge_model_bytes_per_launch": m["alg_bytes"],
                     "merge_model_rate_gbs": m["alg_bytes"] / (kms * 1e-3) / 1e9},
        "counters": counters,
        "phase_cycles_frac": {k: round(v / tot, 4) for k, v in prof.items() if k not in ("total", "waves")},
        "cpu_baseline": cpu,
        "serving": serving,
        **legs,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
