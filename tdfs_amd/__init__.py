"""tdfs_amd -- T-DFS subgraph matching on MI355X.

Python front end over libtdfs.so (include/tdfs.h).  The names mirror the
reference's host interface (STMatch::GraphPreprocessor, src/graph.h:26-213;
STMatch::PatternPreprocessor, src/pattern.h:65-720; cu_test.cu's launch and
reduction): load or generate a CSR graph, plan a pattern, upload the graph
once, count.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import Plan, Result, Tuning, check, i32p, i64p, lib, TdfsError

__all__ = [
    "CSRGraph", "Plan", "Result", "Tuning", "TdfsError", "Device", "DeviceGraph", "Batch",
    "plan_pattern", "load_pattern", "reorder_pattern", "gen_pattern", "estimate_work",
]


@dataclass
class CSRGraph:
    """Symmetric CSR: rowptr int64[n+1], colidx int32[nnz] (each list strictly
    ascending), labels int32[n] or None."""

    rowptr: np.ndarray
    colidx: np.ndarray
    labels: np.ndarray | None = None

    @property
    def n(self) -> int:
        return int(self.rowptr.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.colidx.shape[0])

    @property
    def dmax(self) -> int:
        return int(np.diff(self.rowptr).max()) if self.n else 0

    def degrees(self) -> np.ndarray:
        return np.diff(self.rowptr)

    def validate(self) -> None:
        check(lib().tdfs_graph_validate(self.n, self.nnz, i64p(self.rowptr), i32p(self.colidx)), "validate")

    # --- reference binary format (src/graph.h:136-193) ---
    @classmethod
    def read_bin(cls, prefix: str, labeled: bool = False) -> "CSRGraph":
        meta = np.zeros(4, dtype=np.int64)
        check(lib().tdfs_graph_read_meta(prefix.encode(), i64p(meta)), "read meta")
        n, nnz = int(meta[0]), int(meta[1])
        rowptr = np.zeros(n + 1, dtype=np.int64)
        colidx = np.zeros(max(nnz, 1), dtype=np.int32)
        labels = np.zeros(max(n, 1), dtype=np.int32) if labeled else None
        check(lib().tdfs_graph_read(prefix.encode(), n, nnz, i64p(rowptr), i32p(colidx), i32p(labels)), "read graph")
        return cls(rowptr, colidx[:nnz].copy(), labels[:n].copy() if labeled else None)

    def write_bin(self, prefix: str) -> None:
        check(lib().tdfs_graph_write(prefix.encode(), self.n, self.nnz, i64p(self.rowptr), i32p(self.colidx),
                                     i32p(self.labels)), "write graph")

    # --- seeded generators ---
    @classmethod
    def erdos_renyi(cls, n: int, m: int, seed: int) -> "CSRGraph":
        rowptr = np.zeros(n + 1, dtype=np.int64)
        colidx = np.zeros(max(2 * m, 1), dtype=np.int32)
        check(lib().tdfs_gen_er(n, m, seed, i64p(rowptr), i32p(colidx)), "gen ER")
        return cls(rowptr, colidx[: 2 * m].copy())

    @classmethod
    def chung_lu(cls, n: int, m: int, gamma: float, dmax: float, seed: int) -> "CSRGraph":
        rowptr = np.zeros(n + 1, dtype=np.int64)
        colidx = np.zeros(max(2 * m, 1), dtype=np.int32)
        check(lib().tdfs_gen_chung_lu(n, m, gamma, dmax, seed, i64p(rowptr), i32p(colidx)), "gen Chung-Lu")
        return cls(rowptr, colidx[: 2 * m].copy())

    @classmethod
    def from_edges(cls, n: int, edges) -> "CSRGraph":
        """Builds a symmetric CSR from undirected edges (duplicates and loops dropped)."""
        adj = [set() for _ in range(n)]
        for a, b in edges:
            if a != b:
                adj[a].add(b)
                adj[b].add(a)
        rowptr = np.zeros(n + 1, dtype=np.int64)
        cols = []
        for v in range(n):
            nb = sorted(adj[v])
            cols.extend(nb)
            rowptr[v + 1] = rowptr[v] + len(nb)
        return cls(rowptr, np.asarray(cols, dtype=np.int32))

    def relabeled(self, perm: np.ndarray) -> "CSRGraph":
        """The same graph with vertex v renamed perm[v] (lists re-sorted)."""
        n = self.n
        perm = np.asarray(perm, dtype=np.int64)
        inv = np.empty(n, dtype=np.int64)
        inv[perm] = np.arange(n)
        deg = np.diff(self.rowptr)
        rp = np.zeros(n + 1, dtype=np.int64)
        rp[1:] = np.cumsum(deg[inv])
        # one sort of (new row, new column) keys packed in 64 bits
        key = perm[np.repeat(np.arange(n), deg)] << 32
        key |= perm[self.colidx]
        key.sort(kind="stable")
        labels = None if self.labels is None else np.ascontiguousarray(self.labels[inv]).astype(np.int32)
        return CSRGraph(rp, (key & 0xFFFFFFFF).astype(np.int32), labels)

    def with_labels(self, nlabels: int, seed: int) -> "CSRGraph":
        labels = np.zeros(max(self.n, 1), dtype=np.int32)
        check(lib().tdfs_gen_labels(self.n, nlabels, seed, i32p(labels)), "gen labels")
        return CSRGraph(self.rowptr, self.colidx, labels[: self.n].copy())


def estimate_work(g: CSRGraph, plan: Plan, nsample: int = 2000, seed: int = 1, cap: float = 1e8) -> dict:
    """Sampled DFS work estimate on a host CSR (tdfs_estimate_work)."""
    out = (C.c_double * 3)()
    check(lib().tdfs_estimate_work(g.n, g.nnz, i64p(g.rowptr), i32p(g.colidx), i32p(g.labels if plan.labeled else None),
                                   C.byref(plan), nsample, seed, cap, out), "estimate")
    return {"cost": out[0], "raw": out[1], "capped": bool(out[2])}


def plan_pattern(text: str, labeled: bool) -> Plan:
    """Parses .g text and builds its plan (PatternPreprocessor, src/pattern.h:90-101)."""
    p = Plan()
    check(lib().tdfs_pattern_parse(text.encode(), int(labeled), C.byref(p)), "pattern parse")
    return p


def counting_order(plan: Plan) -> tuple[Plan, list[int]]:
    """The plan the counting kernel runs for `plan` (free_order tuning on) and
    the pattern vertex at each of its positions."""
    q = Plan()
    order = np.zeros(max(plan.k, 1), dtype=np.int32)
    check(lib().tdfs_plan_counting_order(C.byref(plan), C.byref(q), i32p(order)), "counting order")
    return q, [int(x) for x in order[: plan.k]]


def load_pattern(path: str, labeled: bool) -> Plan:
    p = Plan()
    check(lib().tdfs_pattern_load(str(path).encode(), int(labeled), C.byref(p)), "pattern load")
    return p


def reorder_pattern(text: str) -> tuple[str, list[int]]:
    buf = C.create_string_buffer(4096)
    order = (C.c_int32 * 7)()
    check(lib().tdfs_pattern_reorder(text.encode(), order, buf, 4096), "pattern reorder")
    out = buf.value.decode()
    k = sum(1 for l in out.splitlines() if l.startswith("v"))
    return out, list(order)[:k]


def gen_pattern(k: int, m: int, nlabels: int, seed: int) -> str:
    buf = C.create_string_buffer(4096)
    check(lib().tdfs_gen_pattern(k, m, nlabels, seed, buf, 4096), "gen pattern")
    return buf.value.decode()


class DeviceGraph:
    def __init__(self, dev: "Device", g: CSRGraph):
        self.dev = dev
        self.host = g
        self.handle = C.c_void_p()
        check(lib().tdfs_graph_upload(dev.handle, g.n, g.nnz, i64p(g.rowptr), i32p(g.colidx), i32p(g.labels),
                                      C.byref(self.handle)), "graph upload")

    def info(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        check(lib().tdfs_graph_info(self.handle, i64p(out)), "graph info")
        return {"n": int(out[0]), "nnz": int(out[1]), "ordered": bool(out[2]), "hubs": int(out[3])}

    def estimate(self, plan: Plan, nsample: int = 2000, seed: int = 1, cap: float = 1e8) -> dict:
        """Work estimate on the numbering the device counts this plan on."""
        out = (C.c_double * 3)()
        check(lib().tdfs_graph_estimate(self.handle, C.byref(plan), nsample, seed, cap, out), "graph estimate")
        return {"cost": out[0], "raw": out[1], "capped": bool(out[2])}

    def order(self) -> np.ndarray | None:
        """New id of each vertex in the degree-ordered copy, or None."""
        if not self.info()["ordered"]:
            return None
        perm = np.zeros(max(self.host.n, 1), dtype=np.int32)
        check(lib().tdfs_graph_order(self.handle, i32p(perm)), "graph order")
        return perm[: self.host.n]

    def free(self) -> None:
        if self.handle:
            lib().tdfs_graph_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Plans uploaded once; the launch sequence captured into a hipGraph on the
    batch's own stream (tdfs_batch_create) and replayed on demand."""

    def __init__(self, dev: "Device", dg: DeviceGraph, plans: list[Plan], blocks_per_cu: int = 0):
        self.n = len(plans)
        self._plans = (Plan * self.n)(*plans)
        self.handle = C.c_void_p()
        check(lib().tdfs_batch_create(dev.handle, dg.handle, self._plans, self.n, blocks_per_cu,
                                      C.byref(self.handle)), "batch create")

    def launch(self) -> None:
        check(lib().tdfs_batch_launch(self.handle), "batch launch")

    def done(self) -> bool:
        rc = lib().tdfs_batch_done(self.handle)
        if rc < 0:
            check(rc, "batch done")
        return rc == 1

    def wait(self) -> list[Result]:
        res = (Result * self.n)()
        check(lib().tdfs_batch_wait(self.handle, res), "batch wait")
        return list(res)

    def run(self) -> list[Result]:
        self.launch()
        return self.wait()

    def close(self) -> None:
        if self.handle:
            lib().tdfs_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Device:
    """A libtdfs context: one HIP device, one stream, launch workspace."""

    def __init__(self, device: int = 0):
        self.handle = C.c_void_p()
        check(lib().tdfs_ctx_create(device, C.byref(self.handle)), "ctx create")

    def info(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        check(lib().tdfs_ctx_info(self.handle, i64p(out)), "ctx info")
        return {"cus": int(out[0]), "wallclock_khz": int(out[1]), "waves": int(out[2]), "blocks_per_cu": int(out[3])}

    def last_profile(self) -> dict:
        """Shader-clock cycles of the last stats launch, summed over waves."""
        out = (C.c_uint64 * 8)()
        check(lib().tdfs_ctx_last_profile(self.handle, out), "last profile")
        keys = ["filter", "extend", "leaf", "idle", "rebuild", "donate", "total", "waves"]
        return {k: int(v) for k, v in zip(keys, out)}

    def last_counters(self) -> dict:
        """Operation counters of the last stats launch, summed over waves."""
        out = (C.c_uint64 * 8)()
        check(lib().tdfs_ctx_last_counters(self.handle, out), "last counters")
        keys = ["leaf_batches", "leaf_owners", "leaf_units", "leaf_setup_cycles", "extensions", "probe_elements",
                "search_tests", "memo_calls"]
        return {k: int(v) for k, v in zip(keys, out)}

    def tuning(self) -> Tuning:
        t = Tuning()
        check(lib().tdfs_ctx_get_tuning(self.handle, C.byref(t)), "get tuning")
        return t

    def set_tuning(self, **kw) -> None:
        t = self.tuning()
        names = {f[0] for f in type(t)._fields_}
        for k, v in kw.items():
            if k not in names:
                raise TypeError(f"unknown tuning field {k!r}")
            setattr(t, k, v)
        check(lib().tdfs_ctx_set_tuning(self.handle, C.byref(t)), "set tuning")

    def set_stats(self, on: bool) -> None:
        check(lib().tdfs_ctx_set_stats(self.handle, int(on)), "set stats")

    def upload(self, g: CSRGraph) -> DeviceGraph:
        return DeviceGraph(self, g)

    def count(self, dg: DeviceGraph, plan: Plan, edge_begin: int = 0, edge_end: int = -1) -> Result:
        r = Result()
        check(lib().tdfs_count(self.handle, dg.handle, C.byref(plan), edge_begin, edge_end, C.byref(r)), "count")
        return r

    def count_batch(self, dg: DeviceGraph, plans: list[Plan]) -> list[Result]:
        arr = (Plan * len(plans))(*plans)
        res = (Result * len(plans))()
        check(lib().tdfs_count_batch(self.handle, dg.handle, arr, len(plans), res), "count batch")
        return list(res)

    def enumerate(self, dg: DeviceGraph, plan: Plan, cap: int, edge_begin: int = 0,
                  edge_end: int = -1) -> tuple[np.ndarray, Result]:
        """Embeddings reached by the counting DFS (tdfs_enumerate): an int32
        array of min(raw, cap) rows of k vertex ids (plan order, caller's
        numbering, unspecified row order) and the launch's Result."""
        k = int(plan.k)
        rows = np.zeros((max(cap, 1), k), dtype=np.int32)
        written = C.c_uint64()
        r = Result()
        check(lib().tdfs_enumerate(self.handle, dg.handle, C.byref(plan), edge_begin, edge_end,
                                   rows.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(written), C.byref(r)),
              "enumerate")
        return rows[: int(written.value)], r

    def close(self) -> None:
        if self.handle:
            lib().tdfs_ctx_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
