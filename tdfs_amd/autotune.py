"""Autotuner for the batched launch.

Coordinate descent over the launch knobs of tdfs_tuning -- register budget
(waves_per_simd: as compiled without spills, or 4 / 5 / 6 / 8 with), persistent-grid
occupancy (blocks_per_cu: waves resident per CU, i.e. how many DFS stacks
share a CU's registers and LDS), level-0 chunk (chunk_edges: edges per cursor
grab), donation threshold (timeout_us), donated piece size (split_chunk) and
leaf batch size (leaf_pair_cap), the smallest level that gets a hoisted
intersection (hoist_min), intra-workgroup stealing (steal) and the smallest
degree given a neighbour bitmap
(hub_min_degree, read at upload: the graph is uploaded again) --
timing the hipGraph-replayed batch (median of `reps` replays per point; every
replay's counts must equal the statistics launch).  The winner per workload is
written to a JSON file that bench.py reads.

    python -m tdfs_amd.autotune --workload c4 --out profiles/autotune.json
"""
from __future__ import annotations

import argparse
import json
import statistics
import sys
import time
from pathlib import Path

SPACE = {
    "waves_per_simd": [0, 4, 5, 6, 8],
    "blocks_per_cu": [2, 3, 4],
    "chunk_edges": [16, 32, 64, 128],
    "timeout_us": [25, 100, 200, 1000],
    "split_chunk": [1, 4],
    "leaf_pair_cap": [1024, 4096, 16384],
    "hub_min_degree": [64, 128, 512],
    "hoist_min": [1, 8, 32],
    "steal": [1, 2],
}
# knobs read when the graph is uploaded: a new value means a new upload
UPLOAD_KNOBS = ("hub_min_degree",)


class Uploads:
    """The graph uploaded under the current values of the upload knobs."""

    def __init__(self, dev, g):
        self.dev, self.g, self.key, self.dg = dev, g, None, None

    def get(self, knobs: dict):
        key = tuple(knobs.get(k) for k in UPLOAD_KNOBS)
        if key != self.key:
            if self.dg is not None:
                self.dg.free()
            self.dev.set_tuning(**knobs)
            self.dg = self.dev.upload(self.g)
            self.key = key
        return self.dg

    def free(self):
        if self.dg is not None:
            self.dg.free()
            self.dg = None


def time_point(dev, up: Uploads, plans, counts, knobs: dict, reps: int) -> float:
    import tdfs_amd as T

    dg = up.get(knobs)
    dev.set_tuning(**knobs)
    b = T.Batch(dev, dg, plans)
    ts = []
    try:
        for _ in range(reps):
            res = b.run()
            assert [int(r.count) for r in res] == counts, knobs
            ts.append(res[0].kernel_ms)
    finally:
        b.close()
    return statistics.median(ts)


DEFAULT = {"waves_per_simd": 0, "blocks_per_cu": 0, "chunk_edges": 64, "timeout_us": 200, "split_chunk": 1,
           "leaf_pair_cap": 4096, "hub_min_degree": 256, "hoist_min": 2, "steal": 2}


def tune(dev, g, plans, reps: int = 3, rounds: int = 2, log=print, start: dict | None = None,
         knobs: list[str] | None = None) -> dict:
    """Coordinate descent from `start` (default: the library defaults) over
    `knobs` (default: every knob of SPACE); `g` is the host graph (CSRGraph),
    uploaded again whenever an upload knob changes."""
    dev.set_stats(False)
    best = dict(DEFAULT, **(start or {}))
    up = Uploads(dev, g)
    counts = [int(r.count) for r in dev.count_batch(up.get(best), plans)]
    space = {k: v for k, v in SPACE.items() if knobs is None or k in knobs}
    best_ms = time_point(dev, up, plans, counts, best, reps)
    trials = [{"knobs": dict(best), "ms": best_ms}]
    log(f"start {best} {best_ms:.2f} ms")
    for _ in range(rounds):
        improved = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best, **{knob: v})
                ms = time_point(dev, up, plans, counts, cand, reps)
                trials.append({"knobs": cand, "ms": ms})
                log(f"{cand} {ms:.2f} ms")
                if ms < best_ms * 0.98:
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    up.free()
    return {"best": best, "best_ms": best_ms, "trials": trials}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--out", default="profiles/autotune.json")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--rounds", type=int, default=1)
    ap.add_argument("--knobs", default="", help="comma-separated subset of knobs to sweep (default all)")
    ap.add_argument("--resume", action="store_true", help="start from the best point already in --out")
    ap.add_argument("--trials-out", default="", help="also write every sweep trial (knobs, ms) to this file")
    args = ap.parse_args()
    root = Path(__file__).resolve().parent.parent
    sys.path.insert(0, str(root))
    import bench
    import tdfs_amd as T

    g, pats, desc, _ = bench.load_workload(args.workload)
    dev = T.Device(0)
    plans = [T.plan_pattern(t, lab) for t, lab in pats.values()]
    out = Path(args.out)
    data = json.loads(out.read_text()) if out.exists() else {}
    start = data.get(args.workload, {}).get("best") if args.resume else None
    knobs = [k for k in args.knobs.split(",") if k] or None
    t0 = time.time()
    res = tune(dev, g, plans, reps=args.reps, rounds=args.rounds, log=lambda s: print(s, flush=True),
               start=start, knobs=knobs)
    res["workload"] = desc
    res["seconds"] = round(time.time() - t0, 1)
    # the record keeps the winner; the sweep log goes next to it only on request
    trials = res.pop("trials")
    if args.trials_out:
        Path(args.trials_out).write_text(json.dumps(trials, indent=1))
    data[args.workload] = res
    out.write_text(json.dumps(data, indent=1))
    print(json.dumps({"workload": args.workload, "best": res["best"], "best_ms": res["best_ms"]}))
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
