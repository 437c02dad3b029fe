"""Writes tests/golden/nx_counts.json: monomorphism counts computed by
networkx 3.4.2 (GraphMatcher.subgraph_monomorphisms_iter) for small seeded
graphs and patterns.  networkx is independent of both the oracle and the
product; the committed JSON is the fixture.

    python tests/golden/make_nx_fixtures.py
"""
from __future__ import annotations

import json
from pathlib import Path

import networkx as nx
import numpy as np
from networkx.algorithms import isomorphism

HERE = Path(__file__).resolve().parent


def parse_g(text: str):
    labels, edges = [], []
    for line in text.splitlines():
        if line.startswith("v"):
            _, i, l = line.split()
            labels.append(int(l))
        elif line.startswith("e"):
            _, a, b, _l = line.split()
            edges.append((int(a), int(b)))
    return labels, edges


def to_g(k, labels, edges):
    s = "".join(f"v {i} {labels[i]}\n" for i in range(k))
    s += "".join(f"e {a} {b} 1\n" for a, b in edges)
    return s


def random_pattern(rng, k):
    """Connected pattern in an order where every u >= 1 has an earlier
    neighbour and vertex 2 is adjacent to vertex 1."""
    while True:
        adj = np.zeros((k, k), dtype=int)
        for u in range(1, k):
            adj[u, rng.integers(0, u)] = 1
        extra = rng.integers(0, k * (k - 1) // 2 - (k - 1) + 1)
        for _ in range(extra):
            a, b = rng.choice(k, 2, replace=False)
            adj[a, b] = 1
        adj = np.maximum(adj, adj.T)
        np.fill_diagonal(adj, 0)
        if adj[2, 1]:
            edges = [(a, b) for a in range(k) for b in range(a + 1, k) if adj[a, b]]
            return edges


def mono_count(G, P, labeled):
    nm = (lambda a, b: a["l"] == b["l"]) if labeled else None
    gm = isomorphism.GraphMatcher(G, P, node_match=nm)
    return sum(1 for _ in gm.subgraph_monomorphisms_iter())


def main():
    rng = np.random.default_rng(20240601)
    graphs = []
    specs = [("er18", nx.gnp_random_graph(18, 0.5, seed=11)),
             ("ba20", nx.barabasi_albert_graph(20, 4, seed=12)),
             ("hub15", nx.gnp_random_graph(15, 0.25, seed=13))]
    # give hub15 a hub adjacent to everything
    for v in range(1, 15):
        specs[2][1].add_edge(0, v)
    pats = {}
    pdir = HERE / "patterns"
    for i in range(1, 15):
        pats[f"ref{i}"] = (pdir / f"{i}.g").read_text()
    for j in range(12):
        k = int(rng.integers(3, 7))
        edges = random_pattern(rng, k)
        labels = [int(x) for x in rng.integers(0, 2, size=k)]
        pats[f"rnd{j}"] = to_g(k, labels, edges)
    for name, G in specs:
        n = G.number_of_nodes()
        labels = [int(x) for x in rng.integers(0, 4, size=n)]
        for v in G.nodes:
            G.nodes[v]["l"] = labels[v]
        entry = {"name": name, "n": n, "edges": sorted([list(map(int, e)) for e in G.edges]), "labels": labels,
                 "counts": {}}
        for pname, text in pats.items():
            plabels, pedges = parse_g(text)
            k = len(plabels)
            for labeled in (0, 1):
                P = nx.Graph()
                for i in range(k):
                    P.add_node(i, l=plabels[i] if labeled else 0)
                P.add_edges_from(pedges)
                if not labeled:
                    for v in G.nodes:
                        G.nodes[v]["l0"] = 0
                entry["counts"][f"{pname}:{labeled}"] = mono_count(G, P, labeled)
        graphs.append(entry)
    out = {"generator": "networkx " + nx.__version__, "patterns": pats, "graphs": graphs}
    (HERE / "nx_counts.json").write_text(json.dumps(out, indent=1, sort_keys=True))
    print("wrote", HERE / "nx_counts.json")


if __name__ == "__main__":
    main()
