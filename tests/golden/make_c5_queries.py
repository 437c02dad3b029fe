"""Writes tests/golden/c5_queries.json: the seeded c5 query stream (200
queries, seed 5, 30% labeled with 8 labels, budget 2e10) drawn by
tdfs_amd.serve.query_stream against the c5 graph in the numbering the device
counts on (vertices renumbered by ascending degree).

    python tests/golden/make_c5_queries.py
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import tdfs_amd as T  # noqa: E402
from tdfs_amd import serve, workloads as W  # noqa: E402


class HostEstimator:
    def __init__(self, g):
        self.g = g

    def estimate(self, plan, nsample, seed, cap):
        return T.estimate_work(self.g, plan, nsample, seed, cap)


def main():
    G = W.graph("c5")
    D = G.relabeled(W.degree_order(G))
    qs = serve.query_stream(HostEstimator(D), 200, 5, 0.3, 8, 2e10, progress=True)
    out = {"rule": "tdfs_amd/serve.py::query_stream(count=200, seed=5, labeled_frac=0.3, nlabels=8, budget=2e10)",
           "queries": qs}
    (HERE / "c5_queries.json").write_text(json.dumps(out, indent=1))
    print("wrote", HERE / "c5_queries.json")


if __name__ == "__main__":
    main()
