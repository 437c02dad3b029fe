"""GPU parity at the benchmark configurations' full sizes.

Every configuration graph is generated at its real size and uploaded whole.
Each pattern (and each c5 query) is counted on seeded level-0 edge windows
[lo, hi) of the numbering the device counts on (the degree-ordered copy),
and each window's raw count, algorithmic bytes and extension count must equal
the oracle's count_ranges on the same windows of the same numbering
(cu_test.cu:100-105: raw = sum of the per-warp counts, x |Aut| for ulb).
Window widths are chosen by the device's own statistics launch so that the
oracle's work per window stays bounded (the width halves until the window's
algorithmic bytes fit the budget); one window per pattern sits at the end of
a hub row, where the heaviest subtrees start.

c4's patterns are also counted together through the serving API (one launch,
per-plan windows), and c5's queries are answered by tdfs_serve_* both on
windows (against the oracle) and over the whole graph (against tdfs_count
alone).
"""
import json
import os
import time
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
import tdfs_amd as T
from tdfs_amd import serve, workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
THREADS = max(1, min(16, os.cpu_count() or 1))
_T0 = time.time()


def say(*a):
    line = " ".join(str(x) for x in a)
    print(f"[gpu-case +{time.time() - _T0:7.1f}s]", line, flush=True)
    log = os.environ.get("TDFS_CASE_LOG")
    if log:
        with open(log, "a") as f:
            f.write(f"+{time.time() - _T0:.1f}s {line}\n")


@pytest.fixture(scope="module")
def dev():
    d = T.Device(0)
    yield d
    d.close()


def _prepared(dev, G):
    dg = dev.upload(G)
    H = G.relabeled(dg.order())
    return dg, H, O.Prepared(H.rowptr, H.colidx, H.labels)


@pytest.fixture(scope="module")
def cl11(dev):
    """The c2/c3/c4 graph: Chung-Lu n=1.1M m=3.0M gamma=2.1 seed 2, labels
    U{0..3} seed 3 (c2 counts its unlabeled patterns on the same CSR)."""
    say("generate + upload Chung-Lu n=1.1M m=3.0M")
    G = W.graph("c3")
    dg, H, prep = _prepared(dev, G)
    yield G, dg, H, prep
    prep.close()
    dg.free()


@pytest.fixture(scope="module")
def cl4(dev):
    """The c5 graph: Chung-Lu n=4M m=34M gamma=2.3 seed 5, labels U{0..7}."""
    say("generate + upload Chung-Lu n=4M m=34M")
    G = W.graph("c5")
    dg, H, prep = _prepared(dev, G)
    yield G, dg, H, prep
    prep.close()
    dg.free()


def hub_window(H, width, rank=50):
    """The last `width` edges of the row of the vertex `rank` places below the
    top of the degree order: edges from a hub to higher-degree hubs."""
    r = H.n - 1 - rank
    hi = int(H.rowptr[r + 1])
    return max(int(H.rowptr[r]), hi - width), hi


def windows(dev, dg, H, plan, rng, nrand, width, budget):
    """Seeded windows: up to nrand random windows holding at least one match
    (at most 4 nrand draws; empty ones are kept only when no draw holds a
    match) and one hub window, each shrunk (width halved) until its
    statistics launch reports at most `budget` algorithmic bytes.  Returns
    [(lo, hi, stats result)]."""
    assert plan.adj[2][1] or plan.deg[2] <= plan.deg[0], "counted on the degree-ordered copy"
    dev.set_stats(True)

    def fit(lo, hi):
        while True:
            r = dev.count(dg, plan, lo, hi)
            if r.alg_bytes <= budget or hi - lo <= 1:
                return hi, r
            hi = lo + (hi - lo) // 2

    out, empty = [], []
    for lo in rng.integers(0, H.nnz - width, size=4 * nrand):
        if len(out) == nrand:
            break
        lo = int(lo)
        hi, r = fit(lo, lo + width)
        if r.alg_bytes > 20 * budget:
            say("  window", lo, hi, "skipped: one edge holds", r.alg_bytes, "algorithmic bytes")
        elif r.raw > 0:
            out.append((lo, hi, r))
        elif len(empty) < nrand:
            empty.append((lo, hi, r))
    out += empty[: nrand - len(out)]
    lo, hi = hub_window(H, width)
    hi, r = fit(lo, hi)
    if r.alg_bytes <= 20 * budget:
        out.append((lo, hi, r))
    dev.set_stats(False)
    return out


def check_windows(dev, dg, prep, name, text, labeled, wins, **tune):
    """Each window: statistics launch and timed launch (tuned) against the
    oracle on the same window."""
    plan = T.plan_pattern(text, labeled)
    q = O.plan(text, labeled)
    total = 0
    for lo, hi, rs in wins:
        say(f"  {name} window [{lo}, {hi})")
        want = prep.count_ranges(q, [lo], [hi], threads=THREADS)
        assert rs.raw == want["raw"], (name, lo, hi)
        assert rs.alg_bytes == want["alg_bytes"], (name, lo, hi)
        assert rs.extensions == want["extensions"], (name, lo, hi)
        base = dev.tuning()
        dev.set_tuning(**tune)
        r = dev.count(dg, plan, lo, hi)
        dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_})
        assert r.raw == want["raw"] and r.count == want["count"], (name, lo, hi, tune)
        total += want["raw"]
    return total


def test_c2_patterns_full_graph_windows(dev, cl11):
    """c2: 4-clique, diamond and 5-cycle, unlabeled."""
    G, dg, H, prep = cl11
    rng = np.random.default_rng(202)
    for name, (text, lab) in W.patterns("c2", full=True).items():
        say("c2", name)
        plan = T.plan_pattern(text, lab)
        wins = windows(dev, dg, H, plan, rng, 4, 256, 1.5e9)
        raw = check_windows(dev, dg, prep, name, text, lab, wins)
        check_windows(dev, dg, prep, name, text, lab, wins[:2], waves_per_simd=5, timeout_us=50)
        say("c2", name, "raw over windows", raw)


def test_c3_patterns_full_graph_windows(dev, cl11):
    """c3: labeled house, 5-clique, 6-path and 6-cycle."""
    G, dg, H, prep = cl11
    rng = np.random.default_rng(303)
    for name, (text, lab) in W.patterns("c3", full=True).items():
        say("c3", name)
        plan = T.plan_pattern(text, lab)
        wins = windows(dev, dg, H, plan, rng, 4, 256, 1.5e9)
        raw = check_windows(dev, dg, prep, name, text, lab, wins)
        check_windows(dev, dg, prep, name, text, lab, wins[:2], waves_per_simd=5, timeout_us=50)
        say("c3", name, "raw over windows", raw)


def test_c4_patterns_full_graph_windows(dev, cl11):
    """c4: every committed pattern on two windows with tdfs_count, then all
    of them at once through one serving launch with per-plan windows."""
    G, dg, H, prep = cl11
    pats = json.loads((GOLDEN / "c4_patterns.json").read_text())["patterns"]
    rng = np.random.default_rng(404)
    jobs = []
    for p in pats:
        name = f"p{p['index']:02d}_k{p['k']}{'L' if p['labeled'] else ''}"
        say("c4", name)
        plan = T.plan_pattern(p["text"], p["labeled"])
        wins = windows(dev, dg, H, plan, rng, 1, 128, 4e8)
        check_windows(dev, dg, prep, name, p["text"], p["labeled"], wins, waves_per_simd=5)
        jobs += [(plan, lo, hi, rs.raw) for lo, hi, rs in wins]
    say("c4 serving launch with", len(jobs), "windowed plans")
    sv = serve.Server(dev, dg, slots=1, blocks_per_cu=2, max_batch=len(jobs))
    tickets = {sv.submit(plan, lo, hi): want for plan, lo, hi, want in jobs}
    got = {}
    t0 = time.time()
    while len(got) < len(tickets):
        for tk, r, _lat in sv.poll():
            assert not r.flags
            got[tk] = r.raw
        assert time.time() - t0 < 120
    sv.close()
    assert got == tickets


def test_c5_queries_served_full_graph(dev, cl4):
    """c5: 24 committed queries through tdfs_serve_* -- on seeded windows
    against the oracle, and over the whole graph against tdfs_count alone."""
    G, dg, H, prep = cl4
    qs = json.loads((GOLDEN / "c5_queries.json").read_text())["queries"][:24]
    rng = np.random.default_rng(505)
    jobs = []
    for q in qs:
        plan = T.plan_pattern(q["text"], q["labeled"])
        say("c5 query", q["index"], "k", q["k"], "labeled" if q["labeled"] else "unlabeled")
        wins = windows(dev, dg, H, plan, rng, 1, 256, 4e8)
        oq = O.plan(q["text"], q["labeled"])
        for lo, hi, rs in wins:
            want = prep.count_ranges(oq, [lo], [hi], threads=THREADS)
            assert rs.raw == want["raw"] and rs.alg_bytes == want["alg_bytes"], (q["index"], lo, hi)
            jobs.append((plan, lo, hi, want["raw"]))
    sv = serve.Server(dev, dg, slots=4, blocks_per_cu=1, max_batch=8)
    say("c5 serving", len(jobs), "windows")
    tickets = {sv.submit(plan, lo, hi): want for plan, lo, hi, want in jobs}
    got = {}
    t0 = time.time()
    while len(got) < len(tickets):
        for tk, r, _lat in sv.poll():
            assert not r.flags
            got[tk] = r.raw
        assert time.time() - t0 < 120
    assert got == tickets
    # whole graph: served (mixed with each other) == alone
    plans = [T.plan_pattern(q["text"], q["labeled"]) for q in qs]
    alone = []
    for i, p in enumerate(plans):
        say("c5 alone", qs[i]["index"])
        alone.append(dev.count(dg, p).count)
    say("c5 serving whole graph")
    r = serve.run_stream(sv, plans, 1000.0, 55)
    sv.close()
    assert not r["incomplete"] and not r["failed"]
    assert r["counts"] == alone
