"""Serving front end: a continuous-batching query server over a resident graph.

`Server` wraps tdfs_serve_* (include/tdfs.h): queries are submitted as plans,
wait in a FIFO, and start in groups of up to `max_batch` on whichever of the
`slots` concurrent launches (own stream and task ring, grid capped to
`blocks_per_cu` workgroups per CU) is idle, once a region of the server's
DFS-stack arena (`stack_mb`) large enough for the group's largest pattern is
free.  Each query is collected as soon as its own work has ended (per-plan
completion records in host-mapped memory), not when its launch ends.

`run_stream` replays a seeded Poisson arrival process against a server and
reports p50/p99 latency (submit -> result), throughput and matches/s.

    python -m tdfs_amd.serve --rates 2,5,50,1000 --queries 200 \
        --queries-file tests/golden/c5_queries.json --time-limit-ms 3000
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import DeviceGraph, Device, Plan, Result, gen_pattern, plan_pattern
from ._lib import TdfsError, check, i64p, lib


class ServeStartError(TdfsError):
    """tdfs_serve_poll could not start a launch (its queries stay queued and
    start on a later poll); `results` holds the queries that the same poll
    did collect."""

    def __init__(self, code: int, results: list):
        super().__init__(code, "serve poll")
        self.results = results


class Server:
    def __init__(self, dev: Device, dg: DeviceGraph, slots: int = 4, blocks_per_cu: int = 1, max_batch: int = 32,
                 budget: float = 0.0, stack_mb: int = 0):
        self.handle = C.c_void_p()
        check(lib().tdfs_serve_create(dev.handle, dg.handle, slots, blocks_per_cu, max_batch, stack_mb, budget,
                                      C.byref(self.handle)), "serve create")
        self.cap = slots * max_batch

    def submit(self, plan: Plan, edge_begin: int = 0, edge_end: int = -1) -> int:
        """Queues a query (optionally over level-0 edges [edge_begin, edge_end)
        of the numbering the device counts the plan on); returns its ticket."""
        t = C.c_uint64()
        check(lib().tdfs_serve_submit_range(self.handle, C.byref(plan), edge_begin, edge_end, C.byref(t)),
              "serve submit")
        return int(t.value)

    def poll(self) -> list[tuple[int, Result, float]]:
        tickets = (C.c_uint64 * self.cap)()
        res = (Result * self.cap)()
        lat = (C.c_double * self.cap)()
        n = C.c_int32()
        rc = lib().tdfs_serve_poll(self.handle, tickets, res, lat, self.cap, C.byref(n))
        out = [(int(tickets[i]), res[i], float(lat[i])) for i in range(n.value)]
        if rc != 0:
            raise ServeStartError(rc, out)
        return out

    def status(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        check(lib().tdfs_serve_status(self.handle, i64p(out)), "serve status")
        return {"waiting": int(out[0]), "busy_slots": int(out[1]), "slots": int(out[2]), "running": int(out[3])}

    def close(self) -> None:
        if self.handle:
            lib().tdfs_serve_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def query_stream(dg, count: int, seed: int, labeled_frac: float, nlabels: int, budget: float,
                 nsample: int = 4000, max_attempts: int = 50, progress: bool = False) -> list[dict]:
    """Seeded queries: k uniform in 3..7, edge count uniform in [k-1, k(k-1)/2],
    labeled with probability labeled_frac (nlabels labels); a draw whose
    sampled work estimate is capped or above budget is redrawn: after every
    max_attempts rejections with one vertex fewer, and once at 3 vertices
    with labels.  `dg` is a DeviceGraph or anything with the same
    estimate(plan, nsample, seed, cap) method."""
    out = []
    for i in range(count):
        rng0 = np.random.default_rng([seed, i])
        k = int(rng0.integers(3, 8))
        labeled = bool(rng0.random() < labeled_frac)
        attempt = 0
        while True:
            if attempt > 0 and attempt % max_attempts == 0:
                if k > 3:
                    k -= 1
                else:
                    labeled = True
            rng = np.random.default_rng([seed, i, attempt + 1])
            m = int(rng.integers(k - 1, k * (k - 1) // 2 + 1))
            text = gen_pattern(k, m, nlabels if labeled else 0, int(rng.integers(1 << 62)))
            plan = plan_pattern(text, labeled)
            est = dg.estimate(plan, nsample, 1, budget)
            if not est["capped"] and est["cost"] <= budget:
                out.append({"index": i, "k": k, "m": m, "labeled": labeled, "text": text, "est_cost": est["cost"],
                            "attempts": attempt + 1})
                break
            attempt += 1
        if progress and (i + 1) % 25 == 0:
            print(f"query_stream: {i + 1}/{count} queries drawn", file=sys.stderr, flush=True)
    return out


def run_stream(server: Server, plans: list[Plan], rate: float, seed: int, progress: int = 0) -> dict:
    """Submits plans at seeded Poisson arrival times (rate queries/s), polls
    until all are answered; latency = submit -> collected.  Queries flagged
    incomplete (time limit) or failed are listed and left out of matches and
    matches/s (their counts are not exact)."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(plans)))
    results = {}
    lat = {}
    t0 = time.perf_counter()
    nxt = 0
    ticket_of = {}
    start_errors = 0
    while len(results) < len(plans):
        now = time.perf_counter() - t0
        while nxt < len(plans) and arrivals[nxt] <= now:
            ticket_of[server.submit(plans[nxt])] = nxt
            nxt += 1
        try:
            got = server.poll()
        except ServeStartError as e:  # a launch could not start: its queries stay queued
            got = e.results
            start_errors += 1
            if start_errors > 100:
                raise
        for tk, r, l in got:
            i = ticket_of[tk]
            results[i] = r
            lat[i] = l
            if progress and len(results) % progress == 0:
                print(f"  rate {rate:g}: {len(results)}/{len(plans)} answered, "
                      f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if nxt < len(plans) and not server.status()["running"]:
            time.sleep(max(0.0, min(0.001, arrivals[nxt] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    lats = np.array([lat[i] for i in range(len(plans))])
    counts = [int(results[i].count) for i in range(len(plans))]
    incomplete = [i for i in range(len(plans)) if results[i].flags & 1]
    failed = [i for i in range(len(plans)) if results[i].flags & 2]
    exact = [c for i, c in enumerate(counts) if not results[i].flags & 3]
    return {
        "incomplete": incomplete,
        "failed": failed,
        "rate_qps": rate,
        "queries": len(plans),
        "wall_s": wall,
        "throughput_qps": len(plans) / wall,
        "p50_ms": float(np.percentile(lats, 50) * 1e3),
        "p99_ms": float(np.percentile(lats, 99) * 1e3),
        "mean_ms": float(lats.mean() * 1e3),
        "matches": int(sum(exact)),
        "matches_per_s": sum(exact) / wall,
        "start_errors": start_errors,
        "counts": counts,
    }


def main() -> int:
    from . import workloads as W

    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="2,5,50,1000")
    ap.add_argument("--queries", type=int, default=200)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--blocks-per-cu", type=int, default=1)
    ap.add_argument("--waves-per-simd", type=int, default=0,
                    help="register budget of the match kernel (tdfs_tuning.waves_per_simd: 0, 4, 5, 6 or 8)")
    ap.add_argument("--max-batch", type=int, default=8)
    ap.add_argument("--budget", type=float, default=2e10)
    ap.add_argument("--check", type=int, default=20, help="queries recounted alone for equality")
    ap.add_argument("--queries-file", default="", help="committed query list (json) instead of drawing one")
    ap.add_argument("--time-limit-ms", type=int, default=3000,
                    help="per-launch limit: work still pending is dropped and its queries flagged incomplete")
    args = ap.parse_args()
    t = time.time()
    g = W.graph("c5")
    print(f"graph {g.n} vertices {g.nnz} directed edges, {time.time() - t:.1f} s", file=sys.stderr, flush=True)
    dev = Device(0)
    dg = dev.upload(g)
    print(f"uploaded {time.time() - t:.1f} s", file=sys.stderr, flush=True)
    if args.queries_file:
        from pathlib import Path
        qs = json.loads(Path(args.queries_file).read_text())["queries"][: args.queries]
    else:
        qs = query_stream(dg, args.queries, 5, 0.3, 8, args.budget, progress=True)
    plans = [plan_pattern(q["text"], q["labeled"]) for q in qs]
    print(f"{len(plans)} queries {time.time() - t:.1f} s", file=sys.stderr, flush=True)
    out = {"workload": "c5: Chung-Lu n=4M m=34M gamma=2.3 seed 5, 8 labels; seeded query stream (k 3-7, 30% labeled)",
           "budget": args.budget, "slots": args.slots, "blocks_per_cu": args.blocks_per_cu,
           "max_batch": args.max_batch, "waves_per_simd": args.waves_per_simd, "streams": []}
    dev.set_tuning(time_limit_ms=args.time_limit_ms, waves_per_simd=args.waves_per_simd)
    server = Server(dev, dg, args.slots, args.blocks_per_cu, args.max_batch, args.budget)
    alone = {}
    for i in range(min(args.check, len(plans))):
        r = dev.count(dg, plans[i])
        alone[i] = int(r.count)
        print(f"alone {i}: k={plans[i].k} labeled={plans[i].labeled} count={r.count} {r.kernel_ms:.2f} ms",
              file=sys.stderr, flush=True)
    for rate in [float(x) for x in args.rates.split(",")]:
        r = run_stream(server, plans, rate, 50 + int(rate), progress=25)
        same = all(r["counts"][i] == c for i, c in alone.items() if i not in r["incomplete"])
        r["alone_equal"] = same
        del r["counts"]
        out["streams"].append(r)
        print(json.dumps(r), file=sys.stderr, flush=True)
    server.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
