"""GPU parity: libtdfs on cuda:0 against the oracle, bit-exact.

Every case prints its name before touching the GPU so that a failing run
names it."""
import json
import math
import os
import time
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
import tdfs_amd as T
from tdfs_amd.workloads import with_pattern_labels

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


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


def fixture(i):
    return (GOLDEN / "patterns" / f"{i}.g").read_text()


_relabeled = {}


def device_numbering(dg, G, plan):
    """The CSR the device counted on: the degree-ordered copy when the graph
    has one and the plan's count cannot depend on the numbering."""
    perm = dg.order()
    sound = plan.adj[2][1] or plan.deg[2] <= plan.deg[0]
    if perm is None or not sound:
        return G
    key = (id(dg), id(G))
    if key not in _relabeled:
        _relabeled[key] = (dg, G, G.relabeled(perm))  # dg, G kept alive so that their ids stay theirs
    return _relabeled[key][2]


_oracle = {}


def oracle_counts(G, H, text, labeled):
    """The oracle on the caller's numbering and on the device's (kept per
    graph and pattern: several modes of a test compare against them)."""
    key = (id(G), id(H), text, labeled)
    if key not in _oracle:
        q = O.plan(text, labeled)
        want = O.count(G.rowptr, G.colidx, G.labels if labeled else None, q, threads=4)
        want_dev = want if H is G else O.count(H.rowptr, H.colidx, H.labels if labeled else None, q, threads=4)
        assert want_dev["count"] == want["count"]  # the count does not depend on the numbering
        _oracle[key] = (G, H, want, want_dev)  # G, H kept alive so that their ids stay theirs
    return _oracle[key][2], _oracle[key][3]


def check(dev, dg, G, text, labeled, stats=True, **tune):
    p = T.plan_pattern(text, labeled)
    H = device_numbering(dg, G, p)
    want, want_dev = oracle_counts(G, H, text, labeled)
    dev.set_stats(stats)
    if tune:
        dev.set_tuning(**tune)
    r = dev.count(dg, p)
    assert r.raw == want["raw"], (text, labeled, tune)
    assert r.count == want["count"]
    if stats:
        assert r.alg_bytes == want_dev["alg_bytes"], (text, labeled, tune)
        assert r.extensions == want_dev["extensions"], (text, labeled, tune)
        # the timed kernel (the one measured) on the same case
        dev.set_stats(False)
        r2 = dev.count(dg, p)
        assert r2.raw == want["raw"], (text, labeled, tune, "timed kernel")
    return r


def test_smoke_triangle(dev):
    say("triangle ER(2000, 10000)")
    G = T.CSRGraph.erdos_renyi(2000, 10000, 1)
    dg = dev.upload(G)
    check(dev, dg, G, "v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 0 2 1\ne 1 2 1\n", 0)


def test_fixtures_small_graphs(dev):
    graphs = {
        "er": T.CSRGraph.erdos_renyi(3000, 24000, 5).with_labels(4, 1),
        "cl": T.CSRGraph.chung_lu(20000, 80000, 2.1, 1200, 6).with_labels(4, 2),
    }
    for gname, G in graphs.items():
        dg = dev.upload(G)
        for i in range(1, 15):
            for labeled in (0, 1):
                say(gname, f"{i}.g", "labeled" if labeled else "unlabeled")
                check(dev, dg, G, fixture(i), labeled)
        dg.free()


def test_networkx_fixtures(dev):
    d = json.loads((GOLDEN / "nx_counts.json").read_text())
    for g in d["graphs"]:
        G = T.CSRGraph.from_edges(g["n"], g["edges"])
        G.labels = np.asarray(g["labels"], dtype=np.int32)
        dg = dev.upload(G)
        for key, want in g["counts"].items():
            pname, labeled = key.split(":")
            say("nx", g["name"], key)
            p = T.plan_pattern(d["patterns"][pname], int(labeled))
            assert dev.count(dg, p).count == want, (g["name"], key)
        dg.free()


def test_random_patterns(dev):
    G = T.CSRGraph.chung_lu(8000, 40000, 2.2, 500, 11).with_labels(3, 4)
    dg = dev.upload(G)
    rng = np.random.default_rng(12)
    for j in range(40):
        k = int(rng.integers(3, 8))
        m = int(rng.integers(max(k - 1, k * (k - 1) // 2 - 4), k * (k - 1) // 2 + 1))
        text = T.gen_pattern(k, m, 3, int(rng.integers(1 << 62)))
        for labeled in (0, 1):
            say("random", j, k, m, labeled)
            check(dev, dg, G, text, labeled)


def test_sparse_random_patterns(dev):
    """Sparse seeded patterns (trees and near-trees, k 4-6), whose hoisted
    intersections mostly do not involve the previous vertex; default and
    forced hoisting, with donation."""
    G = T.CSRGraph.chung_lu(1500, 6000, 2.1, 60, 23).with_labels(3, 6)
    dg = dev.upload(G)
    base_t = dev.tuning()
    rng = np.random.default_rng(24)
    for j in range(20):
        k = int(rng.integers(4, 7))
        m = int(rng.integers(k - 1, k + 2))
        text = T.gen_pattern(k, m, 3, int(rng.integers(1 << 62)))
        for labeled in (0, 1):
            mode = dict(hoist_min=1, timeout_us=1, donate_levels=5) if j % 2 else dict()
            say("sparse", j, k, m, labeled, mode)
            check(dev, dg, G, text, labeled, **mode)
            dev.set_tuning(**{f: getattr(base_t, f) for f, _ in base_t._fields_ if f != "reserved"})
    dg.free()


def free_levels(plan):
    """Vertices w (2 <= w <= k-3) that no later vertex has as a backward
    neighbour or shares a label with (labeled plans)."""
    k = plan.k
    return [w for w in range(2, k - 2)
            if plan.labeled and all(plan.label[v] != plan.label[w]
                                    and w not in [plan.bn[v][i] for i in range(plan.num_bn[v])]
                                    for v in range(w + 1, k))]


def test_free_levels_timed_kernel(dev):
    """Labeled patterns with free middle vertices: the timed kernel counts
    one subtree per free level and multiplies; counts equal the oracle's,
    also with donation forced at every check (work donated from inside a
    measured subtree carries the level's multiplier to its receiver, and a
    received task is never passed on before it is worked on), with one-edge
    chunks and with the 5-wave kernel."""
    G = T.CSRGraph.chung_lu(6000, 30000, 2.1, 600, 91).with_labels(4, 92)
    dg = dev.upload(G)
    base = dev.tuning()
    rng = np.random.default_rng(93)
    cases = []
    for j in range(60):
        k = int(rng.integers(5, 8))
        m = int(rng.integers(k - 1, k + 3))
        text = T.gen_pattern(k, m, 4, int(rng.integers(1 << 62)))
        p = T.plan_pattern(text, 1)
        if m >= k and free_levels(p):  # no trees: their counts keep the oracle busy for minutes
            cases.append((text, p))
        if len(cases) == 8:
            break
    assert len(cases) == 8
    dev.set_stats(False)
    for j, (text, p) in enumerate(cases):
        want = O.count(G.rowptr, G.colidx, G.labels, O.plan(text, 1), threads=4)["raw"]
        for mode in (dict(), dict(timeout_us=1, donate_levels=5), dict(chunk_edges=1, timeout_us=1),
                     dict(waves_per_simd=5, timeout_us=1, donate_levels=5)):
            say("free levels", j, free_levels(p), mode)
            dev.set_tuning(**mode)
            assert dev.count(dg, p).raw == want, (j, mode)
            dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_})
    dg.free()


def test_tails_timed_kernel(dev):
    """Patterns whose last vertices share their backward neighbours, are
    pairwise non-adjacent and are totally ordered (unlabeled) or share a
    label (labeled): the timed kernel counts their completions as C(n, t) or
    n!/(n-t)! of the first one's candidates; counts equal the oracle's,
    with donation forced too."""
    G = T.CSRGraph.chung_lu(5000, 25000, 2.1, 400, 95).with_labels(2, 96)
    dg = dev.upload(G)
    base = dev.tuning()
    star4 = "v 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\n"
    book3 = "v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 0 4 1\ne 1 2 1\ne 1 3 1\ne 1 4 1\n"
    k23 = "v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\ne 0 1 1\ne 2 1 1\ne 0 3 1\ne 2 3 1\ne 0 4 1\ne 2 4 1\n"
    tri_tail = "v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\ne 0 1 1\ne 0 2 1\ne 1 2 1\ne 2 3 1\ne 2 4 1\n"
    # labeled tails of several labels: leaves 2, 3 (labels 1, 0; 3 shares
    # vertex 1's label); leaves 2, 3, 4 (labels 0, 1, 0; 2 and 4 share the
    # centre's label)
    star4_mixed = "v 0 1\nv 1 0\nv 2 1\nv 3 0\ne 0 1 1\ne 0 2 1\ne 0 3 1\n"
    star5_mixed = "v 0 0\nv 1 1\nv 2 0\nv 3 1\nv 4 0\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 0 4 1\n"
    # a tail (4, 5 over 0, 1) below a vertex (3, over 0 and 2) it must differ
    # from: counted over 3's candidates with one exclusion (labeled: 3 shares
    # the tail's label; unlabeled: NE only)
    above = "e 0 1 1\ne 0 2 1\ne 1 2 1\ne 0 3 1\ne 2 3 1\ne 0 4 1\ne 1 4 1\ne 0 5 1\ne 1 5 1\n"
    tail_excl = "v 0 1\nv 1 1\nv 2 0\nv 3 1\nv 4 1\nv 5 1\n" + above
    dev.set_stats(False)
    for name, text in (("star4", star4), ("book3", book3), ("k23", k23), ("triangle+2 pendants", tri_tail),
                       ("star4 mixed labels", star4_mixed), ("star5 mixed labels", star5_mixed),
                       ("tail below a vertex it must differ from", tail_excl)):
        for labeled in (1,) if "mixed" in name else (0, 1):
            p = T.plan_pattern(text, labeled)
            want = O.count(G.rowptr, G.colidx, G.labels if labeled else None, O.plan(text, labeled), threads=4)["raw"]
            for mode in (dict(), dict(timeout_us=1, donate_levels=5)):
                say("tail", name, labeled, mode)
                dev.set_tuning(**mode)
                assert dev.count(dg, p).raw == want, (name, labeled, mode)
                dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_})
    dg.free()


def test_free_order_timed_kernel(dev):
    """Labeled patterns of the 64-pattern set that the counting kernel
    matches in another order (free levels moved forward, tails of several
    labels moved to the end):
    counts equal the oracle's with the reorder on and off, with donation
    forced and with the 5-wave kernel; the statistics kernel keeps the plan's
    order (its algorithmic bytes equal the oracle's)."""
    G = T.CSRGraph.chung_lu(6000, 30000, 2.1, 600, 99).with_labels(4, 100)
    dg = dev.upload(G)
    base = dev.tuning()
    pats = json.loads((GOLDEN / "c4_patterns.json").read_text())["patterns"]
    cases = [e for e in pats if e["index"] in (3, 7, 9, 23, 29, 38, 43, 44, 53, 58, 63)]
    assert len(cases) == 11 and all(e["labeled"] for e in cases)
    for e in cases:
        check(dev, dg, G, e["text"], 1)  # statistics and timed kernels
        want = oracle_counts(G, device_numbering(dg, G, T.plan_pattern(e["text"], 1)), e["text"], 1)[0]["raw"]
        dev.set_stats(False)
        for mode in (dict(free_order=2), dict(timeout_us=1, donate_levels=5), dict(waves_per_simd=5, chunk_edges=1)):
            say("free order", e["index"], mode)
            dev.set_tuning(**mode)
            assert dev.count(dg, T.plan_pattern(e["text"], 1)).raw == want, (e["index"], mode)
            dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_})
    dg.free()


def leaf_factor(plan):
    """The level after which the last vertex's completions no longer depend
    on the path (its backward neighbours fixed, every later vertex it must
    differ from always among its candidates, no LT / GT on it), or -1."""
    k, v = plan.k, plan.k - 1
    if k < 5:
        return -1
    bn = [plan.bn[v][i] for i in range(plan.num_bn[v])]
    a = max(max(bn), 1)
    if a > k - 4 or (not plan.labeled and any(plan.cond[v][j] != 2 for j in range(v))):
        return -1
    for j in range(a + 1, v):
        if (not plan.labeled or plan.label[j] == plan.label[v]) and not all(plan.adj[j][b] for b in bn):
            return -1
    return a


def test_leaf_factor_timed_kernel(dev):
    """Patterns whose last vertex depends on an early prefix only (a
    pendant, a vertex over the first edge, labeled ones over later
    vertices): the counting kernel drops it and multiplies each prefix's
    count by its completions; counts equal the oracle's, also with donation
    forced at every check (donated edges and subtrees carry the factor), with
    one-edge chunks, with the 5-wave kernel and with the factor off.  Every
    vertex has a neighbour bitmap, so closing-triangle shortcuts are taken
    wherever they apply (never across the vertex the factor is taken at)."""
    G = T.CSRGraph.chung_lu(6000, 30000, 2.1, 600, 99).with_labels(4, 100)
    base = dev.tuning()
    dev.set_tuning(hub_min_degree=1, hub_budget_mb=4096)
    dg = dev.upload(G)
    dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_})
    pats = {e["index"]: e for e in json.loads((GOLDEN / "c4_patterns.json").read_text())["patterns"]}
    cases = [(f"p{i:02d}", pats[i]["text"], pats[i]["labeled"]) for i in (2, 19, 22, 34, 39, 49, 57)]
    k5_apex = "".join(f"v {i} {l}\n" for i, l in enumerate([3, 0, 2, 1, 0, 1])) + "".join(
        f"e {a} {b} 1\n" for a, b in [(i, j) for i in range(5) for j in range(i + 1, 5)] + [(0, 5), (1, 5), (2, 5)])
    cases.append(("K5 + apex over its first triangle (labeled)", k5_apex, True))
    cases.append(("K4 + pendant on an edge vertex",
                  "v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 1 2 1\ne 1 3 1\ne 2 3 1\n"
                  "e 0 4 1\n", False))
    assert all(leaf_factor(T.counting_order(T.plan_pattern(t, l))[0]) >= 1 for _, t, l in cases)
    dev.set_stats(False)
    for name, text, labeled in cases:
        want = O.count(G.rowptr, G.colidx, G.labels if labeled else None, O.plan(text, labeled), threads=4)["raw"]
        for mode in (dict(), dict(timeout_us=1, donate_levels=5), dict(chunk_edges=1, timeout_us=1),
                     dict(waves_per_simd=5), dict(tails=2)):
            say("leaf factor", name, mode)
            dev.set_tuning(**mode)
            assert dev.count(dg, T.plan_pattern(text, labeled)).raw == want, (name, mode)
            dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_})
    dg.free()


def triangle_tail(plan):
    """Unlabeled plans whose last three vertices a, b, c form a triangle over
    one set of earlier neighbours, are totally ordered among themselves
    (no cycle) and each stand to every earlier vertex as a's own condition
    on it says (closure of the LT / GT conditions)."""
    k = plan.k
    if plan.labeled or k < 5:
        return False
    a, b, c = k - 3, k - 2, k - 1
    bn = lambda v: {plan.bn[v][i] for i in range(plan.num_bn[v])}
    if bn(b) != bn(a) | {a} or bn(c) != bn(a) | {a, b}:
        return False
    less = set()
    for v in range(k):
        for j in range(v):
            if plan.cond[v][j] == 0:  # TDFS_LT: M(v) < M(j)
                less.add((v, j))
            if plan.cond[v][j] == 1:  # TDFS_GT
                less.add((j, v))
    for m in range(k):
        for x in range(k):
            for y in range(k):
                if (x, m) in less and (m, y) in less:
                    less.add((x, y))
    if any((x, x) in less for x in (a, b, c)):
        return False
    if not all((x, y) in less or (y, x) in less for x in (a, b, c) for y in (a, b, c) if x < y):
        return False
    return all(((v, j) in less) == (plan.cond[a][j] == 0) and ((j, v) in less) == (plan.cond[a][j] == 1)
               for j in range(a) for v in (a, b, c))


def test_triangle_tail_timed_kernel(dev):
    """Cliques K5-K7, whose last three vertices form a triangle over all the
    earlier ones: the timed kernel counts a path's completions as the
    triangles of its candidate set (unlabeled) or across the candidates of
    the last three labels (labeled, three distinct labels; some shared with
    earlier vertices) when those are at most 64 vertices (adjacency from
    either vertex's bitmap row, else a search of a list); counts equal the
    oracle's with every vertex a hub, with some, with none, with donation
    forced and with the 5-wave kernel."""
    G = T.CSRGraph.chung_lu(3000, 24000, 2.1, 300, 97).with_labels(4, 98)
    base = dev.tuning()
    clique = lambda labs: "".join(f"v {i} {l}\n" for i, l in enumerate(labs)) + "".join(
        f"e {i} {j} 1\n" for i in range(len(labs)) for j in range(i + 1, len(labs)))
    cases = [(f"K{k}", clique([1] * k), 0) for k in (5, 6, 7)]
    assert all(triangle_tail(T.plan_pattern(t, 0)) for _, t, _l in cases)
    # labeled: last three of distinct labels (repeating earlier ones), and a
    # control whose last two share a label (expanded level by level)
    cases += [("K5 labeled", clique([0, 1, 2, 3, 0]), 1), ("K6 labeled", clique([1, 1, 2, 3, 0, 2]), 1),
              ("K5 labeled, last two alike", clique([0, 1, 2, 3, 3]), 1)]
    want = {name: O.count(G.rowptr, G.colidx, G.labels if lab else None, O.plan(text, lab), threads=4)["raw"]
            for name, text, lab in cases}
    dev.set_stats(False)
    for hub in (dict(hub_min_degree=1, hub_budget_mb=4096), dict(hub_min_degree=40, hub_budget_mb=4096),
                dict(hub_budget_mb=-1)):
        dev.set_tuning(**hub)
        dg = dev.upload(G)
        for mode in (dict(), dict(timeout_us=1, donate_levels=5), dict(waves_per_simd=5), dict(tails=2)):
            dev.set_tuning(**mode)
            for name, text, lab in cases:
                say("triangle tail", name, hub, mode)
                assert dev.count(dg, T.plan_pattern(text, lab)).raw == want[name], (name, hub, mode)
            dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_})
            dev.set_tuning(**hub)
        dg.free()
        dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_})


def test_scheduling_modes_are_exact(dev):
    """Forced scheduler paths: donate at every check, a tiny task ring (full
    ring path), odd chunk sizes, reuse off; counts and algorithmic bytes are
    unchanged."""
    G = T.CSRGraph.chung_lu(1500, 6000, 2.1, 150, 21).with_labels(2, 5)
    dg = dev.upload(G)
    base = dev.tuning()
    modes = [
        dict(timeout_us=1, donate_levels=5),
        dict(timeout_us=1, donate_levels=5, queue_log2=8),
        dict(timeout_us=-1),
        dict(chunk_edges=1),
        dict(chunk_edges=1024, timeout_us=1),
        dict(split_chunk=7, timeout_us=1, donate_levels=5),
        dict(reuse=2),
        dict(leaf_batch=2),
        dict(leaf_batch=2, timeout_us=1, donate_levels=5),
        dict(cursor_order=2),
        dict(waves_per_simd=8),
        dict(waves_per_simd=4, timeout_us=1),
        dict(waves_per_simd=5),
        dict(waves_per_simd=6, timeout_us=1),
        dict(leaf_pair_cap=64, timeout_us=1),
        dict(leaf_pair_cap=100000),
        dict(cursor_order=2, timeout_us=1, chunk_edges=16),
        dict(cursor_order=3),
        dict(hoist_min=-1),
        dict(hoist_min=1, timeout_us=1, donate_levels=5),
        dict(hoist_min=1, reuse=2),
        dict(hoist_min=1, leaf_batch=2),
        dict(hoist_min=64),
        dict(steal=1),
        dict(steal=1, timeout_us=1, donate_levels=5),
        dict(steal=1, waves_per_simd=5),
        dict(steal=1, leaf_batch=2),
        dict(steal=1, hoist_min=1, timeout_us=1),
        dict(steal=1, chunk_edges=1024, timeout_us=-1),
        dict(steal=1, reuse=2, chunk_edges=512),
        dict(merge_min=1),
        dict(merge_min=8, waves_per_simd=5),
        dict(merge_min=1, hoist_min=-1, leaf_batch=2),
        dict(merge_min=4, timeout_us=1, donate_levels=5),
        dict(leaf_memo=2),
        dict(hoist_min=1, timeout_us=1, donate_levels=5, steal=1),
        dict(leaf_memo=2, reuse=2),
        dict(leaf_hash=2),
        dict(leaf_hash=1, leaf_pair_cap=64, timeout_us=1, steal=1),
        dict(leaf_hash=3),
        dict(leaf_hash=3, leaf_pair_cap=64, timeout_us=1),
        dict(ext_hash=2),
        dict(tails=2),
        dict(free_order=2),
        dict(ext_hash=2, leaf_hash=2, leaf_memo=2),
    ]
    for i in (1, 2, 3, 5, 7, 11, 13):
        for labeled in (0, 1):
            for mode in modes:
                say("sched", f"{i}.g", labeled, mode)
                r = check(dev, dg, G, fixture(i), labeled, stats=mode.get("reuse") != 2, **mode)
                dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_ if f != "reserved"})


def test_stealing_takes_work_and_is_exact(dev):
    """With ring donation off and 1,024-edge chunks, idle waves can only get
    work from sibling waves of their workgroup: some work is stolen, and
    counts, algorithmic bytes and extensions still equal the oracle's (the
    thief copies the sibling's lists, it recomputes nothing)."""
    G = T.CSRGraph.chung_lu(20000, 100000, 2.1, 2000, 77).with_labels(3, 7)
    dg = dev.upload(G)
    base_t = dev.tuning()
    stolen = 0
    for i in (1, 2, 3, 7, 11, 13):
        for labeled in (0, 1):
            for mode in (dict(steal=1, timeout_us=-1, chunk_edges=1024),
                         dict(steal=1, timeout_us=-1, chunk_edges=1024, waves_per_simd=5),
                         dict(steal=1, chunk_edges=1024, leaf_batch=2, hoist_min=1)):
                say("steal", f"{i}.g", labeled, mode)
                r = check(dev, dg, G, fixture(i), labeled, stats=mode.get("waves_per_simd") is None, **mode)
                stolen += r.stolen
                dev.set_tuning(**{f: getattr(base_t, f) for f, _ in base_t._fields_})
    assert stolen > 0
    dg.free()


def test_many_labels_without_segments(dev):
    """More than 64 distinct labels: no label segments are built and every
    candidate is label-tested; counts and statistics still match the oracle,
    with and without hoisted intersections."""
    base = T.CSRGraph.chung_lu(4000, 24000, 2.1, 400, 41)
    rng = np.random.default_rng(41)
    labels = rng.integers(0, 2, size=base.n).astype(np.int32)
    rare = rng.choice(base.n, size=200, replace=False)
    labels[rare] = rng.integers(2, 100, size=rare.shape[0])
    G = T.CSRGraph(base.rowptr, base.colidx, labels)
    dg = dev.upload(G)
    base_t = dev.tuning()
    for i in (1, 2, 5, 7, 11, 13):
        text = fixture(i)
        k = sum(1 for l in text.splitlines() if l.startswith("v"))
        labs = np.random.default_rng([41, i]).integers(0, 2, size=k)
        lt = with_pattern_labels(text, labs)
        for mode in (dict(), dict(hoist_min=-1), dict(hoist_min=1, timeout_us=1, donate_levels=5)):
            say("many labels", f"{i}.g", mode)
            check(dev, dg, G, lt, 1, **mode)
            dev.set_tuning(**{f: getattr(base_t, f) for f, _ in base_t._fields_ if f != "reserved"})
    dg.free()


def test_dense_labeled_patterns_hoisted(dev):
    """Dense 5-/6-vertex patterns of the c4 set, whose hoisted intersections
    depend on an early prefix only (built once per M(1) or M(2) and shared by
    the levels below), counted with hoisting on, off and forced, with
    donation; counts and statistics equal the oracle's."""
    G = T.CSRGraph.chung_lu(3000, 24000, 2.1, 300, 71).with_labels(4, 7)
    dg = dev.upload(G)
    base_t = dev.tuning()
    pats = json.loads((GOLDEN / "c4_patterns.json").read_text())["patterns"]
    for idx in (3, 23, 29, 33, 49, 13):
        p = pats[idx]
        for labeled in sorted({0, int(p["labeled"])}):
            for mode in (dict(), dict(hoist_min=-1), dict(hoist_min=1, timeout_us=1, donate_levels=5),
                         dict(hoist_min=1, leaf_batch=2)):
                say("dense", idx, labeled, mode)
                check(dev, dg, G, p["text"], labeled, **mode)
                dev.set_tuning(**{f: getattr(base_t, f) for f, _ in base_t._fields_ if f != "reserved"})
    dg.free()


def test_enumeration_matches_oracle(dev):
    """tdfs_enumerate writes exactly the embeddings the oracle's DFS reaches
    on the numbering the device counted on, mapped back to the caller's ids
    (for unlabeled plans the symmetry-broken representatives depend on the
    numbering); also with donation and hoisting forced.  Rows are cut at cap
    while the count stays complete."""
    G = T.CSRGraph.chung_lu(2000, 10000, 2.1, 200, 81).with_labels(3, 8)
    dg = dev.upload(G)
    base_t = dev.tuning()
    for i in (1, 2, 4, 5, 7, 11):
        text = fixture(i)
        for labeled in (0, 1):
            p = T.plan_pattern(text, labeled)
            q = O.plan(text, labeled)
            H = device_numbering(dg, G, p)
            want, total = O.enumerate_embeddings(H.rowptr, H.colidx, H.labels if labeled else None, q)
            if H is not G:
                inv = np.empty(G.n, dtype=np.int64)
                inv[dg.order()] = np.arange(G.n)
                want = inv[want]
            want = sorted(map(tuple, want.tolist()))
            for mode in (dict(), dict(timeout_us=1, donate_levels=5, hoist_min=1)):
                say("enumerate", f"{i}.g", labeled, mode)
                dev.set_tuning(**mode)
                rows, r = dev.enumerate(dg, p, cap=total + 5)
                dev.set_tuning(**{f: getattr(base_t, f) for f, _ in base_t._fields_ if f != "reserved"})
                assert r.raw == total, (i, labeled, mode)
                assert sorted(map(tuple, rows.tolist())) == want, (i, labeled, mode)
            if total > 3:
                rows, r = dev.enumerate(dg, p, cap=3)
                assert rows.shape[0] == 3 and r.raw == total
                assert set(map(tuple, rows.tolist())) <= set(want)
    dg.free()


def test_dense_core_long_shared_lists(dev):
    """A complete core (K_110) inside a sparse graph: shared candidate lists
    of 100+ elements next to owner lists of similar length (flattened pair
    path; also with leaf_pair_cap = 64 and forced hoisting); counts and
    statistics equal the oracle's, and the 4-clique count is at least the
    closed form n!/(n-4)! of the core."""
    core = 110
    edges = [(a, b) for a in range(core) for b in range(a + 1, core)]
    rng = np.random.default_rng(5)
    n = 3000
    for _ in range(6000):
        a, b = rng.integers(core, n, size=2)
        edges.append((int(a), int(b)))
    G = T.CSRGraph.from_edges(n, edges).with_labels(2, 3)
    dg = dev.upload(G)
    base_t = dev.tuning()
    k4 = T.plan_pattern(fixture(2), False)
    r = dev.count(dg, k4)
    assert r.count >= core * (core - 1) * (core - 2) * (core - 3)
    for i in (1, 2, 7):
        for labeled in (0, 1):
            for mode in (dict(), dict(leaf_pair_cap=64, timeout_us=1), dict(hoist_min=1), dict(merge_min=16),
                         dict(merge_min=1, leaf_batch=2)):
                say("dense core", f"{i}.g", labeled, mode)
                check(dev, dg, G, fixture(i), labeled, **mode)
                dev.set_tuning(**{f: getattr(base_t, f) for f, _ in base_t._fields_ if f != "reserved"})
    dg.free()


def test_cli_drivers_match_oracle(dev, tmp_path):
    """The reference-compatible drivers read the binary graph files written
    by tdfs_graph_write and end stdout with "<pattern>\t<ms>\t<count>"
    (cu_test.cu:107): ulb prints raw x |Aut|, lb the label-preserving count,
    both equal to the oracle's."""
    import subprocess
    G = T.CSRGraph.chung_lu(5000, 25000, 2.1, 400, 91).with_labels(3, 9)
    prefix = tmp_path / "g"
    G.write_bin(str(prefix))
    for i in (1, 2, 4, 7):
        pat = GOLDEN / "patterns" / f"{i}.g"
        text = pat.read_text()
        for name, labeled in (("ulb", 0), ("lb", 1)):
            say("cli", name, f"{i}.g")
            exe = Path(T.__file__).resolve().parent / "bin" / name
            r = subprocess.run([str(exe), str(prefix), str(pat)], capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            last = r.stdout.strip().splitlines()[-1].split("\t")
            assert last[0] == str(pat) and float(last[1]) >= 0.0
            want = O.count(G.rowptr, G.colidx, G.labels if labeled else None, O.plan(text, labeled), threads=4)
            assert int(last[2]) == want["count"], (name, i)


def test_hub_beyond_reference_buffer(dev):
    """A hub of degree 6000 > 4096 (the reference's per-level slot size,
    src/config.h:12): lists are never truncated here."""
    n = 6500
    edges = [(0, v) for v in range(1, 6001)]
    rng = np.random.default_rng(3)
    for _ in range(30000):
        a, b = rng.integers(1, n, size=2)
        edges.append((int(a), int(b)))
    G = T.CSRGraph.from_edges(n, edges).with_labels(2, 8)
    dg = dev.upload(G)
    star = "v 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\n"
    path5 = "v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\ne 0 1 1\ne 1 2 1\ne 2 3 1\ne 3 4 1\n"
    # unlabeled stars and paths through a 6,000-degree hub are counted in the
    # billions (minutes for the oracle): those two run labeled only
    for text, labs in ((fixture(1), (0, 1)), (fixture(4), (0, 1)), (star, (1,)), (path5, (1,))):
        for labeled in labs:
            say("hub", text.count("v "), labeled)
            check(dev, dg, G, text, labeled)


def test_hub_bitmaps_all_or_none(dev):
    """Membership through neighbour bitmaps (every vertex a hub) and through
    binary search only (no bitmaps), on the degree-ordered copy and on the
    numbering as given, give identical counts and bytes."""
    G = T.CSRGraph.chung_lu(3000, 15000, 2.1, 400, 51).with_labels(3, 2)
    base = dev.tuning()
    try:
        for hub in (dict(hub_min_degree=1, hub_budget_mb=4096), dict(hub_budget_mb=-1),
                    dict(hub_budget_mb=-1, degree_order=2), dict(hub_min_degree=1, degree_order=2)):
            dev.set_tuning(**hub)
            dg = dev.upload(G)
            for i in (1, 2, 3, 4, 5, 6, 7, 9, 12, 14):
                for labeled in (0, 1):
                    say("hub bitmaps", hub, f"{i}.g", labeled)
                    check(dev, dg, G, fixture(i), labeled)
            dg.free()
    finally:
        dev.set_tuning(hub_min_degree=base.hub_min_degree, hub_budget_mb=base.hub_budget_mb,
                       degree_order=base.degree_order)


def test_batch_equals_alone(dev):
    G = T.CSRGraph.chung_lu(20000, 100000, 2.1, 2000, 31).with_labels(3, 9)
    dg = dev.upload(G)
    rng = np.random.default_rng(5)
    plans = []
    for j in range(24):
        # dense patterns: sparse 5-6 vertex trees explode on this hub-heavy graph
        k = int(rng.integers(3, 7))
        m = int(rng.integers(max(k - 1, k * (k - 1) // 2 - 2), k * (k - 1) // 2 + 1))
        text = T.gen_pattern(k, m, 3 if j % 2 else 0, int(rng.integers(1 << 62)))
        plans.append(T.plan_pattern(text, j % 2))
    dev.set_stats(True)
    say("batch of", len(plans))
    batch = dev.count_batch(dg, plans)
    dev.set_tuning(cursor_order=2)
    batch2 = dev.count_batch(dg, plans)
    # stealing between waves of different plans, both cursor orders
    base = dev.tuning()
    say("batch with stealing")
    dev.set_tuning(steal=1, timeout_us=20, chunk_edges=256)
    batch3 = dev.count_batch(dg, plans)
    dev.set_tuning(cursor_order=1)
    batch4 = dev.count_batch(dg, plans)
    say("batch with merge path")
    dev.set_tuning(steal=2, merge_min=8)
    batch5 = dev.count_batch(dg, plans)
    dev.set_tuning(**{f: getattr(base, f) for f, _ in base._fields_})
    say("batch with one cursor for the device, and timed")
    dev.set_tuning(cursor_order=3, merge_min=0, steal=2)
    batch6 = dev.count_batch(dg, plans)
    dev.set_stats(False)
    dev.set_tuning(cursor_order=1)
    timed = dev.count_batch(dg, plans)
    dev.set_stats(True)
    for j, p in enumerate(plans):
        say("alone", j)
        alone = dev.count(dg, p)
        for b in (batch, batch2, batch3, batch4, batch5, batch6):
            assert alone.raw == b[j].raw, j
            assert alone.alg_bytes == b[j].alg_bytes, j
        assert alone.raw == timed[j].raw, j


def test_graph_replays_match_eager(dev):
    """A captured batch replayed several times, with eager launches and host
    synchronisation in between, returns the eager counts on every replay."""
    G = T.CSRGraph.chung_lu(20000, 100000, 2.1, 2000, 33).with_labels(4, 6)
    dg = dev.upload(G)
    texts = [fixture(i) for i in (1, 2, 4, 6, 9, 12)]
    plans = [T.plan_pattern(t, i % 2) for i, t in enumerate(texts)]
    dev.set_stats(False)
    eager = [dev.count(dg, p).count for p in plans]
    b = T.Batch(dev, dg, plans)
    for rep in range(4):
        say("graph replay", rep)
        got = [r.count for r in b.run()]
        assert got == eager, rep
        assert dev.count(dg, plans[rep % len(plans)]).count == eager[rep % len(plans)]
    b.close()


def test_concurrent_batches(dev):
    """Two batches sharing the device (capped grids, own streams) run at the
    same time and each returns its counts alone."""
    G = T.CSRGraph.chung_lu(30000, 150000, 2.1, 3000, 34).with_labels(3, 7)
    dg = dev.upload(G)
    pa = [T.plan_pattern(fixture(i), 0) for i in (1, 2, 3)]
    pb = [T.plan_pattern(fixture(i), 1) for i in (8, 11, 13)]
    want_a = [dev.count(dg, p).count for p in pa]
    want_b = [dev.count(dg, p).count for p in pb]
    ba = T.Batch(dev, dg, pa, blocks_per_cu=1)
    bb = T.Batch(dev, dg, pb, blocks_per_cu=1)
    for rep in range(3):
        say("concurrent batches", rep)
        ba.launch()
        bb.launch()
        assert [r.count for r in bb.wait()] == want_b
        assert [r.count for r in ba.wait()] == want_a
    ba.close()
    bb.close()


def test_serving_mix_equals_alone(dev):
    """Queries streamed through the continuous-batching server (several slots
    running side by side, queries arriving while others run) return exactly
    their alone counts."""
    from tdfs_amd import serve
    G = T.CSRGraph.chung_lu(20000, 100000, 2.2, 1500, 61).with_labels(4, 3)
    dg = dev.upload(G)
    qs = serve.query_stream(dg, 40, 9, 0.4, 4, 5e8)
    plans = [T.plan_pattern(q["text"], q["labeled"]) for q in qs]
    alone = [dev.count(dg, p).count for p in plans]
    sv = serve.Server(dev, dg, slots=3, blocks_per_cu=1, max_batch=8, budget=5e8)
    for rate in (200.0, 5000.0):
        say("serving stream", rate)
        r = serve.run_stream(sv, plans, rate, 7)
        assert r["counts"] == alone, rate
        assert r["p99_ms"] >= r["p50_ms"] > 0
    sv.close()
    base = dev.tuning()
    dev.set_tuning(steal=1, timeout_us=20)  # completion records stay exact when work moves between waves
    sv = serve.Server(dev, dg, slots=2, blocks_per_cu=1, max_batch=8)
    say("serving stream with stealing")
    r = serve.run_stream(sv, plans, 5000.0, 8)
    assert r["counts"] == alone
    sv.close()
    dev.set_tuning(steal=base.steal, timeout_us=base.timeout_us)


def test_serving_error_paths(dev, monkeypatch):
    """A launch that cannot start leaves its query queued at the head (it runs
    on a later poll, exactly), and a launch whose run fails reports its query
    once with TDFS_FAILED and frees the slot; every other query stays exact and
    no ticket is lost or reported twice.  Both failures are forced from the
    host (the kernels run normally): launch 2 fails to start, launch 4's run is
    taken as failed."""
    from tdfs_amd import serve
    G = T.CSRGraph.chung_lu(20000, 100000, 2.2, 1500, 63)
    dg = dev.upload(G)
    plans = [T.plan_pattern(fixture(i), 0) for i in (1, 2, 3, 7, 1, 2)]
    alone = [dev.count(dg, p).count for p in plans]
    monkeypatch.setenv("TDFS_SERVE_FAIL_START", "2")
    monkeypatch.setenv("TDFS_SERVE_FAIL_RUN", "4")
    sv = serve.Server(dev, dg, slots=1, blocks_per_cu=1, max_batch=1)
    ticket_of = {sv.submit(p): i for i, p in enumerate(plans)}
    got, start_errors = {}, 0
    t0 = time.time()
    while len(got) < len(plans) and time.time() - t0 < 60:
        try:
            out = sv.poll()
        except serve.ServeStartError as e:
            assert e.code == -5 and "forced" in str(e)
            out = e.results
            start_errors += 1
        for tk, r, _ in out:
            assert tk in ticket_of and tk not in got
            got[tk] = (int(r.count), int(r.flags))
        time.sleep(0.001)
    sv.close()
    say("serving error paths", start_errors, got)
    assert start_errors == 1 and len(got) == len(plans)
    by_query = {ticket_of[tk]: v for tk, v in got.items()}
    # launches: 1 = q0, 2 = q1 (not started), 3 = q1, 4 = q2 (run failed), 5.. = q3..
    for i, (c, f) in sorted(by_query.items()):
        if i == 2:
            assert f & 2, (i, f)
        else:
            assert f == 0 and c == alone[i], (i, c, f, alone[i])


def test_serving_refuses_over_budget(dev):
    from tdfs_amd import serve
    G = T.CSRGraph.chung_lu(20000, 100000, 2.1, 2000, 62)
    dg = dev.upload(G)
    sv = serve.Server(dev, dg, slots=1, blocks_per_cu=1, max_batch=4, budget=1e3)
    with pytest.raises(T.TdfsError) as e:
        sv.submit(T.plan_pattern(fixture(7), 0))
    assert e.value.code == -7
    sv.close()


def test_time_limit_flags_incomplete(dev):
    """A launch that hits its time limit drops the pending work and flags the
    plans it belonged to; a plan that finished before stays exact."""
    G = T.CSRGraph.chung_lu(200000, 1500000, 2.1, 8000, 71)
    dg = dev.upload(G)
    heavy = T.plan_pattern("v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\ne 0 1 1\ne 1 2 1\ne 2 3 1\ne 3 4 1\n", 0)
    tri = T.plan_pattern("v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 0 2 1\ne 1 2 1\n", 0)
    want = dev.count(dg, tri).count
    base = dev.tuning()
    try:
        dev.set_tuning(time_limit_ms=200)
        say("time limit heavy")
        r = dev.count(dg, heavy)
        assert r.flags & 1
        assert r.kernel_ms < 5000
        r = dev.count(dg, tri)
        assert not (r.flags & 1) and r.count == want
    finally:
        dev.set_tuning(time_limit_ms=base.time_limit_ms)


def test_edge_ranges_partition(dev):
    """Level-0 shards (directed-edge ranges) add up to the whole count."""
    G = T.CSRGraph.chung_lu(50000, 250000, 2.1, 4000, 41)
    dg = dev.upload(G)
    for i in (1, 2, 6):
        p = T.plan_pattern(fixture(i), 0)
        whole = dev.count(dg, p).raw
        cuts = [0, 12345, G.nnz // 2, G.nnz - 1, G.nnz]
        say("ranges", f"{i}.g")
        parts = [dev.count(dg, p, a, b).raw for a, b in zip(cuts, cuts[1:])]
        assert sum(parts) == whole


def test_config1_triangle_er(dev):
    """ER G(n=20000, m=100000) seed 1, triangle: count == trace(A^3)."""
    G = T.CSRGraph.erdos_renyi(20000, 100000, 1)
    import scipy.sparse as sp
    A = sp.csr_matrix((np.ones(G.nnz, dtype=np.int64), G.colidx, G.rowptr), shape=(G.n, G.n))
    tr = int((A @ A).multiply(A).sum())
    dg = dev.upload(G)
    say("config1 triangle")
    r = check(dev, dg, G, "v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 0 2 1\ne 1 2 1\n", 0)
    assert r.count == tr
    assert r.raw * 6 == tr


def test_empty_and_tiny(dev):
    E = T.CSRGraph.from_edges(10, [])
    dg = dev.upload(E)
    say("empty graph")
    assert dev.count(dg, T.plan_pattern(fixture(2), 0)).count == 0
    K3 = T.CSRGraph.from_edges(3, [(0, 1), (1, 2), (0, 2)])
    dg = dev.upload(K3)
    assert dev.count(dg, T.plan_pattern(fixture(2), 0)).count == 0
    assert dev.count(dg, T.plan_pattern("v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 0 2 1\ne 1 2 1\n", 0)).count == 6
    assert dev.count(dg, T.plan_pattern(fixture(2), 0), 2, 2).count == 0


def test_labeled_plan_needs_labels(dev):
    G = T.CSRGraph.erdos_renyi(100, 300, 1)
    dg = dev.upload(G)
    with pytest.raises(T.TdfsError):
        dev.count(dg, T.plan_pattern(fixture(8), 1))


def test_arbitrary_label_values(dev):
    """Graph labels are arbitrary int32 values (negative, above 255, sparse):
    the upload maps them densely and pattern labels go through the same map,
    so counts equal the oracle's on the raw values; a pattern label that no
    vertex carries gives 0."""
    base = T.CSRGraph.chung_lu(6000, 30000, 2.1, 500, 43)
    values = np.array([-7, 5, 300, 1 << 20], dtype=np.int32)
    rng = np.random.default_rng(43)
    G = T.CSRGraph(base.rowptr, base.colidx, values[rng.integers(0, 4, size=base.n)])
    dg = dev.upload(G)
    for i in (1, 2, 5, 11):
        text = fixture(i)
        k = sum(1 for l in text.splitlines() if l.startswith("v"))
        for trial in range(3):
            labs = values[np.random.default_rng([43, i, trial]).integers(0, 4, size=k)]
            say("label values", f"{i}.g", labs.tolist())
            check(dev, dg, G, with_pattern_labels(text, labs), 1)
        absent = with_pattern_labels(text, [77] + [int(values[0])] * (k - 1))
        assert dev.count(dg, T.plan_pattern(absent, 1)).count == 0
    dg.free()


def test_serving_reports_each_query_when_its_work_ends(dev):
    """A light query that shares a launch with a heavy one is collected while
    the heavy one still runs (per-plan completion records), with its exact
    count; the heavy one is collected later."""
    from tdfs_amd import serve
    G = T.CSRGraph.chung_lu(200000, 1500000, 2.1, 8000, 71)
    dg = dev.upload(G)
    heavy = T.plan_pattern("v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\ne 0 1 1\ne 1 2 1\ne 2 3 1\ne 3 4 1\n", 0)
    light = T.plan_pattern("v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 0 2 1\ne 1 2 1\n", 0)
    want = dev.count(dg, light).count
    base = dev.tuning()
    try:
        dev.set_tuning(time_limit_ms=1500)
        sv = serve.Server(dev, dg, slots=1, blocks_per_cu=1, max_batch=2)
        th = sv.submit(heavy)
        tl = sv.submit(light)
        say("serving heavy + light in one launch")
        seen = {}
        t0 = time.time()
        while len(seen) < 2:
            for tk, r, lat in sv.poll():
                seen[tk] = (r, lat, sv.status()["running"])
            assert time.time() - t0 < 60
        r_l, lat_l, running_after_light = seen[tl]
        r_h, lat_h, _ = seen[th]
        assert r_l.count == want and not r_l.flags
        assert lat_l < lat_h
        assert running_after_light >= 1  # the heavy query was still on the device
        sv.close()
    finally:
        dev.set_tuning(time_limit_ms=base.time_limit_ms)


def test_serving_admits_by_stack_memory(dev):
    """A stack arena that holds one 4-vertex launch at a time still answers
    every query exactly (launches wait for memory); a query whose stacks can
    never fit is refused with TDFS_E_BUDGET."""
    from tdfs_amd import serve
    G = T.CSRGraph.chung_lu(20000, 100000, 2.1, 2000, 63).with_labels(3, 4)
    dg = dev.upload(G)
    plans = [T.plan_pattern(fixture(i), i % 2) for i in (1, 2, 3, 5, 8, 11)]
    alone = [dev.count(dg, p).count for p in plans]
    waves = dev.info()["cus"] * 4
    dstride = max(64, (G.dmax + 63) // 64 * 64)
    need4 = waves * (2048 + 2 * dstride) * 4
    sv = serve.Server(dev, dg, slots=3, blocks_per_cu=1, max_batch=2, stack_mb=-(-need4 // (1 << 20)))
    with pytest.raises(T.TdfsError) as e:
        sv.submit(T.plan_pattern(fixture(7), 0))  # 5 vertices: more stack than the arena holds
    assert e.value.code == -7
    four = [j for j, p in enumerate(plans) if p.k <= 4]
    say("serving under a one-launch stack arena", len(four), "queries")
    r = serve.run_stream(sv, [plans[j] for j in four], 500.0, 3)
    assert r["counts"] == [alone[j] for j in four]
    sv.close()
