"""Pins the CPU oracle (oracle/tdfs_oracle.c) before anything is checked
against it: hand-derived plan vectors for the reference's 14 pattern files,
bliss 0.73 built from the reference sources, networkx monomorphism counts,
closed forms, an independent brute-force counter, and invariants."""
import itertools
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
import tdfs_amd as T

GOLDEN = Path(__file__).resolve().parent / "golden"

# |Aut| of data/pattern/1.g..7.g (8..14 are the same shapes):
# diamond 4, K4 24, K4+pendant-to-{0,1} 4, two triangles + vertex 4 ~ {2,3} 4,
# wheel W4 8, K5 - {2,4} 12, K5 120.
AUT = {1: 4, 2: 24, 3: 4, 4: 4, 5: 8, 6: 12, 7: 120}

# Symmetry-breaking pairs derived by hand from the stabiliser-chain rule
# (src/pattern.h:679-707): take the orbit of the smallest vertex with a
# non-trivial orbit, pair its minimum n0 with every other member, fix n0, repeat.
PAIRS = {
    1: [(0, 1), (2, 3)],                     # diamond: {0,1} and {2,3} swap
    2: [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)],  # K4: total order
    3: [(0, 1), (2, 3)],
    4: [(0, 1), (2, 3)],
    5: [(1, 2), (1, 3), (1, 4), (2, 4)],     # wheel: hub 0 fixed; rim D4
    6: [(0, 1), (0, 3), (1, 3), (2, 4)],     # K5-{2,4}: orbits {0,1,3}, {2,4}
    7: [(0, 1), (0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (1, 4), (2, 3), (2, 4), (3, 4)],
}


def fixture(i):
    return (GOLDEN / "patterns" / f"{i}.g").read_text()


@pytest.mark.parametrize("i", range(1, 15))
def test_fixture_plan_vectors(i):
    shape = (i - 1) % 7 + 1
    for labeled in (0, 1):
        p = O.plan(fixture(i), labeled)
        assert p.mult == AUT[shape]
        assert [tuple(p.pairs[j]) for j in range(p.npairs)] == PAIRS[shape]
        # sym01 <=> (0,1) is a pair (vertex 1 in the orbit of 0)
        assert p.sym01 == int((0, 1) in PAIRS[shape])
        # conditions: GT against j iff (j, u) is a pair, else NE; never LT
        for u in range(p.k):
            for j in range(u):
                assert p.cond_op[u][j] == (1 if (j, u) in PAIRS[shape] else 2)


def test_fixture_reuse_levels():
    # diamond: v3 has BN {0,1} = BN(v2) -> reuses C(2); conditions of v2
    # ({GT 0?} no: v2 vs 0 NE, vs 1 NE) must equal those of v3 for j < 2.
    p = O.plan(fixture(1), 0)
    assert list(p.shared_lvl)[:4] == [-1, -1, -1, 2]
    assert p.num_bn_sh[3] == 0
    # wheel: v4 ~ {0,1,3}; v2 ~ {0,1} subset -> dep 2, extra BN {3}
    p = O.plan(fixture(5), 0)
    assert p.shared_lvl[4] == 2 and [p.bn_sh[4][j] for j in range(p.num_bn_sh[4])] == [3]
    # K5 labeled (14.g): labels 0,1,2,3,0 -> v3 cannot reuse v2 (labels differ)
    p = O.plan(fixture(14), 1)
    assert p.shared_lvl[3] == -1
    # K5 unlabeled: v3 reuses v2, v4 reuses v3
    p = O.plan(fixture(7), 0)
    assert list(p.shared_lvl)[:5] == [-1, -1, -1, 2, 3]


def _orbit_chain_pairs(k, group):
    """Closed form of the rule: pairs (n0, o) for every o in the orbit of n0
    under the current stabiliser (independent restatement in Python)."""
    g = list(group)
    pairs = set()
    while True:
        orbits = {}
        for v in range(k):
            orb = {q[v] for q in g}
            orbits.setdefault(min(orb), set()).update(orb)
        big = [r for r in sorted(orbits) if len(orbits[r]) > 1]
        if not big:
            break
        n0 = big[0]
        pairs |= {(n0, o) for o in orbits[n0] if o != n0}
        g = [q for q in g if q[n0] == n0]
    return sorted(pairs)


def _bf_group(k, adj):
    return [q for q in itertools.permutations(range(k))
            if all(adj[i][j] == adj[q[i]][q[j]] for i in range(k) for j in range(k))]


def _random_patterns(n, seed, kmax=7):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(3, kmax + 1))
        m = int(rng.integers(k - 1, k * (k - 1) // 2 + 1))
        nl = int(rng.integers(0, 4))
        out.append(T.gen_pattern(k, m, nl, int(rng.integers(1 << 62))))
    return out


def test_conditions_closed_form_random():
    for text in _random_patterns(150, 1):
        p = O.plan(text, 0)
        adj = [[p.adj[i][j] for j in range(p.k)] for i in range(p.k)]
        grp = _bf_group(p.k, adj)
        assert p.mult == len(grp)
        assert [tuple(p.pairs[j]) for j in range(p.npairs)] == _orbit_chain_pairs(p.k, grp)


def test_bliss_group_matches_bruteforce():
    """bliss 0.73 (compiled from the reference's vendored sources) + closure
    yields the same group as enumeration, so the conditions it feeds are the
    same.  A trivial group gives no generators (empty set) in the reference;
    the oracle treats it as {id}: both give no conditions and |Aut| = 1."""
    texts = [fixture(i) for i in range(1, 8)] + _random_patterns(80, 2)
    if O.bliss_group(3, [[0, 1, 1], [1, 0, 1], [1, 1, 0]]) is None:
        pytest.skip("oracle/_ref/libbliss_ref.so not built (reference sources absent)")
    for text in texts:
        p = O.plan(text, 0)
        adj = [[p.adj[i][j] for j in range(p.k)] for i in range(p.k)]
        bl = O.bliss_group(p.k, adj)
        bf = _bf_group(p.k, adj)
        if len(bf) == 1:
            assert bl == []
        else:
            assert sorted(bl) == sorted(bf)
        assert [tuple(p.pairs[j]) for j in range(p.npairs)] == _orbit_chain_pairs(p.k, bf if bl == [] else bl)


def _csr(n, edges):
    return T.CSRGraph.from_edges(n, edges)


def test_networkx_fixtures():
    d = json.loads((GOLDEN / "nx_counts.json").read_text())
    for g in d["graphs"]:
        G = _csr(g["n"], g["edges"])
        labels = np.asarray(g["labels"], dtype=np.int32)
        for key, want in g["counts"].items():
            pname, labeled = key.split(":")
            labeled = int(labeled)
            p = O.plan(d["patterns"][pname], labeled)
            got = O.count(G.rowptr, G.colidx, labels if labeled else None, p)["count"]
            assert got == want, (g["name"], key)


def test_bruteforce_random_graphs():
    rng = np.random.default_rng(5)
    for trial in range(6):
        n = int(rng.integers(8, 14))
        G = T.CSRGraph.erdos_renyi(n, int(rng.integers(n, n * (n - 1) // 3)), int(rng.integers(1 << 30)))
        lab = G.with_labels(2, trial + 1)
        for text in [fixture(i) for i in range(1, 8)] + _random_patterns(10, trial + 10, kmax=6):
            for labeled in (0, 1):
                p = O.plan(text, labeled)
                L = lab.labels if labeled else None
                want = O.bruteforce(G.rowptr, G.colidx, L, p)
                assert O.count(G.rowptr, G.colidx, L, p)["count"] == want


def test_triangle_trace_closed_form():
    G = T.CSRGraph.erdos_renyi(300, 3000, 9)
    A = np.zeros((G.n, G.n), dtype=np.int64)
    for v in range(G.n):
        A[v, G.colidx[G.rowptr[v]:G.rowptr[v + 1]]] = 1
    tri = O.plan("v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 0 2 1\ne 1 2 1\n", 0)
    assert O.count(G.rowptr, G.colidx, None, tri)["count"] == int(np.trace(A @ A @ A))
    # C4: non-degenerate closed 4-walks = tr(A^4) - 2 sum d^2 + 2m
    c4 = O.plan("v 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 1 2 1\ne 2 3 1\ne 3 0 1\n", 0)
    d = A.sum(1)
    A2 = A @ A
    want = int(np.trace(A2 @ A2)) - 2 * int((d * d).sum()) + int(d.sum())
    assert O.count(G.rowptr, G.colidx, None, c4)["count"] == want


def test_complete_graph_and_star_closed_forms():
    n = 9
    K = _csr(n, [(a, b) for a in range(n) for b in range(a + 1, n)])
    for text in [fixture(i) for i in range(1, 8)] + _random_patterns(20, 3):
        p = O.plan(text, 0)
        assert O.count(K.rowptr, K.colidx, None, p)["count"] == math.perm(n, p.k)
    # stars S_t (centre 0) on a random graph: sum_v d(d-1)..(d-t+1)
    G = T.CSRGraph.erdos_renyi(60, 400, 4)
    d = G.degrees()
    for t in (2, 3, 4, 5, 6):
        # order centre, leaf, leaf... : vertex 2 is not adjacent to vertex 1,
        # but dP(2) = 1 <= dP(0) keeps the level-0 prune sound
        text = "".join(f"v {i} 1\n" for i in range(t + 1)) + "".join(f"e 0 {i} 1\n" for i in range(1, t + 1))
        p = O.plan(text, 0)
        assert p.mult == math.factorial(t)
        want = int(sum(math.perm(int(x), t) for x in d))
        assert O.count(G.rowptr, G.colidx, None, p)["count"] == want


def test_empty_and_small_graphs():
    tri = O.plan(fixture(2), 0)
    E = _csr(5, [])
    assert O.count(E.rowptr, E.colidx, None, tri)["count"] == 0
    P3 = _csr(3, [(0, 1), (1, 2), (0, 2)])  # k > n for K4
    assert O.count(P3.rowptr, P3.colidx, None, tri)["count"] == 0


def test_invariants_reuse_threads_ranges():
    G = T.CSRGraph.chung_lu(3000, 15000, 2.1, 300, 77)
    lab = G.with_labels(3, 5)
    for text in [fixture(i) for i in (1, 2, 3, 5, 6, 7, 10, 12)]:
        for labeled in (0, 1):
            p = O.plan(text, labeled)
            L = lab.labels if labeled else None
            a = O.count(G.rowptr, G.colidx, L, p, threads=1)
            b = O.count(G.rowptr, G.colidx, L, p, threads=4, reuse=False)
            c = O.count(G.rowptr, G.colidx, L, p, threads=3)
            assert a["raw"] == b["raw"] == c["raw"]
            assert a["alg_bytes"] == c["alg_bytes"]
            cuts = [0, G.nnz // 3, G.nnz // 2 + 7, G.nnz]
            parts = [O.count(G.rowptr, G.colidx, L, p, lo, hi)["raw"] for lo, hi in zip(cuts, cuts[1:])]
            assert sum(parts) == a["raw"]


def test_unsound_prune_order_is_reference_semantics():
    """Appendix hazard: for P = {0-1, 0-2, 2-3, 2-4} (vertex 2 not adjacent to
    vertex 1, dP(2) = 3 > dP(0) = 2) the level-0 test 'some neighbour of M(1)
    has degree >= dP(2)' can drop embeddings.  The oracle keeps the
    reference's rule: its count equals brute force restricted to maps whose
    M(1) has such a neighbour."""
    text = "v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\ne 0 1 1\ne 0 2 1\ne 2 3 1\ne 2 4 1\n"
    p = O.plan(text, 0)
    # path 1-0-2 with 2 ~ {3,4}: vertex 1 is a leaf whose only neighbour (0)
    # needs degree >= 2
    G = _csr(7, [(0, 1), (0, 2), (2, 3), (2, 4), (5, 6), (6, 0)])
    d = G.degrees()
    full = O.bruteforce(G.rowptr, G.colidx, None, p)
    # restricted brute force in Python
    n = G.n
    adj = {(a, b) for a in range(n) for b in G.colidx[G.rowptr[a]:G.rowptr[a + 1]]}
    pe = [(0, 1), (0, 2), (2, 3), (2, 4)]
    restricted = 0
    for m in itertools.permutations(range(n), 5):
        if all((m[a], m[b]) in adj for a, b in pe):
            c = m[1]
            if max(d[w] for w in G.colidx[G.rowptr[c]:G.rowptr[c + 1]]) >= 3:
                restricted += 1
    got = O.count(G.rowptr, G.colidx, None, p)["count"]
    assert got == restricted
    assert restricted <= full


def test_counts_invariant_under_renumbering():
    """ulb = #monomorphisms and lb = #label-preserving monomorphisms do not
    depend on the vertex numbering of the data graph (for orders whose
    level-0 prune is sound), while the DFS tree and its bytes do."""
    G = T.CSRGraph.chung_lu(4000, 20000, 2.1, 300, 12).with_labels(3, 4)
    rng = np.random.default_rng(7)
    perm = rng.permutation(G.n)
    H = G.relabeled(perm)
    H.validate()
    deg_order = np.empty(G.n, dtype=np.int64)
    deg_order[np.argsort(G.degrees(), kind="stable")] = np.arange(G.n)
    D = G.relabeled(deg_order)
    for text in [fixture(i) for i in range(1, 15)] + _random_patterns(20, 21, kmax=6):
        for labeled in (0, 1):
            p = O.plan(text, labeled)
            if not (p.adj[2][1] or p.deg[2] <= p.deg[0]):
                continue
            counts = [O.count(X.rowptr, X.colidx, X.labels if labeled else None, p, threads=4)["raw"]
                      for X in (G, H, D)]
            assert counts[0] == counts[1] == counts[2], (text, labeled)


def _all_monomorphisms(G, k, padj, plab, labeled):
    """Every injective map of the pattern into G preserving edges (and labels
    when labeled), by plain backtracking over vertex ids: independent of the
    plan, its conditions and its pruning."""
    nbr = [set(G.colidx[G.rowptr[v]:G.rowptr[v + 1]].tolist()) for v in range(G.n)]
    out = []

    def rec(m):
        d = len(m)
        if d == k:
            out.append(tuple(m))
            return
        for v in range(G.n):
            if v in m or (labeled and G.labels[v] != plab[d]):
                continue
            if all(v in nbr[m[j]] for j in range(d) if padj[d][j]):
                rec(m + [v])
    rec([])
    return out


@pytest.mark.parametrize("i", [1, 2, 3, 4, 5, 8, 9])
def test_enumeration_is_the_counted_set(i):
    """The oracle's enumeration (the counting DFS writing every embedding)
    holds exactly raw distinct embeddings; unlabeled, the images of those
    embeddings under Aut(P) are every monomorphism exactly once (so the
    symmetry-broken set is one representative per automorphism class);
    labeled, they are every label-preserving monomorphism."""
    G = T.CSRGraph.erdos_renyi(24, 110, 30 + i).with_labels(2, i)
    text = fixture(i)
    for labeled in (0, 1):
        p = O.plan(text, labeled)
        k = int(p.k)
        padj = [[int(p.adj[a][b]) for b in range(k)] for a in range(k)]
        plab = [int(p.vlabel[a]) for a in range(k)]
        rows, total = O.enumerate_embeddings(G.rowptr, G.colidx, G.labels if labeled else None, p)
        assert total == O.count(G.rowptr, G.colidx, G.labels if labeled else None, p)["raw"]
        got = [tuple(r) for r in rows.tolist()]
        assert len(got) == len(set(got)) == total
        every = _all_monomorphisms(G, k, padj, plab, labeled)
        if labeled:
            assert sorted(got) == sorted(every)
        else:
            auts = [s for s in itertools.permutations(range(k))
                    if all(padj[s[a]][s[b]] == padj[a][b] for a in range(k) for b in range(k))]
            assert len(auts) == p.mult
            # m o sigma for every representative m and automorphism sigma
            images = [tuple(m[s[a]] for a in range(k)) for m in got for s in auts]
            assert len(images) == len(set(images))
            assert set(images) == set(every)


def test_pins_catch_planted_mistakes():
    """The checks are sensitive: plausible planner / DFS mistakes planted in
    a copy of the oracle's plan (a dropped or flipped symmetry condition, a
    dropped level-0 order filter, a dropped backward neighbour, a wrong
    multiplicity) make the count differ from the brute force, which the
    correct plan matches, or (for count-neutral ones) fail the hand-derived
    plan vectors."""
    import copy
    G = T.CSRGraph.erdos_renyi(30, 150, 77)
    for i in (1, 2, 5, 7):
        good = O.plan(fixture(i), 0)
        want = O.bruteforce(G.rowptr, G.colidx, None, good)
        assert O.count(G.rowptr, G.colidx, None, good)["count"] == want
        k = int(good.k)
        planted = []
        # 1. a GT condition of a DFS level (u >= 2) turned into NE (a
        #    symmetry pair lost); vertex 1's is the level-0 filter (sym01)
        for u in range(2, k):
            for j in range(u):
                if good.cond_op[u][j] == 1:
                    p = copy.deepcopy(good)
                    p.cond_op[u][j] = 2
                    planted.append(("dropped condition", p))
                    break
            else:
                continue
            break
        # 2. a GT condition flipped to LT
        for u in range(2, k):
            for j in range(u):
                if good.cond_op[u][j] == 1:
                    p = copy.deepcopy(good)
                    p.cond_op[u][j] = 0
                    planted.append(("flipped condition", p))
                    break
            else:
                continue
            break
        # 3. the last backward neighbour of the last vertex dropped
        u = k - 1
        if good.num_bn[u] >= 2:
            p = copy.deepcopy(good)
            p.num_bn[u] -= 1
            p.shared_lvl[u] = -1
            p.num_bn_sh[u] = 0
            planted.append(("dropped backward neighbour", p))
        # 3b. the level-0 order filter (r < c) dropped
        if good.sym01:
            p = copy.deepcopy(good)
            p.sym01 = 0
            planted.append(("dropped level-0 order", p))
        # 4. multiplicity off by a factor
        p = copy.deepcopy(good)
        p.mult = good.mult * 2
        planted.append(("wrong multiplicity", p))
        assert planted

        def vectors_ok(p):  # the hand-derived plan vectors of this fixture
            shape = (i - 1) % 7 + 1
            return p.mult == AUT[shape] and p.sym01 == int((0, 1) in PAIRS[shape]) and all(
                p.cond_op[u][j] == (1 if (j, u) in PAIRS[shape] else 2) for u in range(k) for j in range(u))

        assert vectors_ok(good)
        for what, p in planted:
            got = O.count(G.rowptr, G.colidx, None, p, reuse=False)["count"]
            # caught by the count (brute force) or, for count-neutral mistakes
            # such as flipping the order between two interchangeable
            # vertices, by the plan vectors
            assert got != want or not vectors_ok(p), (i, what)
