"""Writes the measured pattern sets of the c4 and c5 benchmarks (run on an
MI355X; the outputs are then committed as tests/golden/c4_patterns.json and
tests/golden/c5_queries.json).

c4: 64 slots, seed 4 (tdfs_amd.workloads.c4_slots: k = 3 + i mod 5, odd i
    labeled with 4 labels), on the Chung-Lu n=1.1M m=3.0M gamma=2.1 graph
    (labels seed 3), gate 500 ms alone with the benchmark's c4 tuning.
c5: 200 queries, seed 5 (workloads.c5_slots: k uniform in 3..7, 30% labeled
    with 8 labels), on the Chung-Lu n=4M m=34M gamma=2.3 graph, gate 100 ms
    alone with the default tuning.

Each slot keeps the first seeded draw that is not isomorphic to a kept
pattern and finishes alone within the gate (workloads.select_patterns); every
attempt and its measured outcome is recorded with the set, so the selection
can be replayed without a GPU.

    python tests/golden/make_pattern_sets.py c4|c5 OUT.json
"""
import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import tdfs_amd as T  # noqa: E402
from tdfs_amd import workloads as W  # noqa: E402

SETS = {
    "c4": dict(graph="c4", slots=W.c4_slots, seed=4, nlabels=4, gate_ms=500, tuned="c4"),
    "c5": dict(graph="c5", slots=W.c5_slots, seed=5, nlabels=8, gate_ms=100, tuned=None),
}


def main() -> int:
    which, out_path = sys.argv[1], Path(sys.argv[2])
    spec = SETS[which]
    t0 = time.time()
    G = W.graph(spec["graph"])
    dev = T.Device(0)
    tuning = {}
    if spec["tuned"]:
        tuning = json.loads((ROOT / "profiles" / "autotune.json").read_text())[spec["tuned"]]["best"]
        dev.set_tuning(**tuning)
    dg = dev.upload(G)
    print(f"graph ready {time.time() - t0:.1f} s", flush=True)
    gate = spec["gate_ms"]
    dev.count(dg, T.plan_pattern(W.TRIANGLE, False))  # warm-up

    def measure(text, labeled):
        dev.set_tuning(time_limit_ms=gate)
        r = dev.count(dg, T.plan_pattern(text, labeled))
        dev.set_tuning(time_limit_ms=0)
        return not (r.flags & 1), r.kernel_ms

    def progress(rec):
        print(f"{time.time() - t0:7.1f}s slot {rec['index']} attempt {rec['attempt']} k={rec['k']} m={rec['m']} "
              f"labeled={rec['labeled']} {rec['outcome']} {rec.get('ms', '')}", flush=True)

    kept, trials = W.select_patterns(spec["slots"](), spec["seed"], spec["nlabels"], measure, progress=progress)
    out = {
        "rule": f"tdfs_amd/workloads.py::select_patterns({which}_slots(), seed={spec['seed']}, "
                f"nlabels={spec['nlabels']}); gate {gate} ms alone on MI355X",
        "gate_ms": gate,
        "tuning": tuning or "defaults",
        ("patterns" if which == "c4" else "queries"): kept,
        # the text of an attempt is kept where it was measured (a replay regenerates the rest)
        "trials": [t if t["outcome"] in ("kept", "over gate") else {f: v for f, v in t.items() if f != "text"}
                   for t in trials],
    }
    out_path.write_text(json.dumps(out, separators=(",", ":")))
    print("wrote", out_path, len(kept), "patterns,", len(trials), "attempts", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
