"""ctypes front end of the CPU oracle (oracle/tdfs_oracle.c) and of the bliss
cross-check library (oracle/_ref/libbliss_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product never
imports this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF = HERE / "_ref" / "libbliss_ref.so"
MAXK = 7


class OrcPlan(C.Structure):
    _fields_ = [
        ("k", C.c_int32),
        ("labeled", C.c_int32),
        ("adj", (C.c_int32 * MAXK) * MAXK),
        ("vlabel", C.c_int32 * MAXK),
        ("deg", C.c_int32 * MAXK),
        ("num_bn", C.c_int32 * MAXK),
        ("bn", (C.c_int32 * MAXK) * MAXK),
        ("cond_op", (C.c_int32 * MAXK) * MAXK),
        ("shared_lvl", C.c_int32 * MAXK),
        ("num_bn_sh", C.c_int32 * MAXK),
        ("bn_sh", (C.c_int32 * MAXK) * MAXK),
        ("sym01", C.c_int32),
        ("npairs", C.c_int32),
        ("pairs", (C.c_int32 * 2) * 64),
        ("mult", C.c_int64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `python -m tdfs_amd.build`")
        L = C.CDLL(str(LIB))
        L.orc_parse_pattern.argtypes = [C.c_char_p, C.c_int, C.POINTER(OrcPlan)]
        L.orc_plan_build.argtypes = [C.POINTER(OrcPlan)]
        L.orc_count.argtypes = [C.c_int32, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                C.POINTER(C.c_int32), C.POINTER(OrcPlan), C.c_int64, C.c_int64, C.c_int, C.c_int,
                                C.POINTER(C.c_uint64)]
        L.orc_count_ranges.argtypes = [C.c_int32, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32), C.POINTER(OrcPlan), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint64)]
        L.orc_prepare.restype = C.c_void_p
        L.orc_prepare.argtypes = [C.c_int32, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32)]
        L.orc_release.argtypes = [C.c_void_p]
        L.orc_enumerate.argtypes = [C.c_int32, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32), C.POINTER(OrcPlan), C.POINTER(C.c_int32), C.c_uint64,
                                    C.POINTER(C.c_uint64)]
        L.orc_count_prepared.argtypes = [C.c_void_p, C.POINTER(OrcPlan), C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint64)]
        L.orc_bruteforce_count.argtypes = [C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                           C.POINTER(C.c_int32), C.POINTER(OrcPlan)]
        L.orc_bruteforce_count.restype = C.c_uint64
        L.orc_plan_sizeof.restype = C.c_size_t
        assert L.orc_plan_sizeof() == C.sizeof(OrcPlan), "OrcPlan layout mismatch"
        _lib = L
    return _lib


def plan(text: str, labeled: bool) -> OrcPlan:
    p = OrcPlan()
    rc = lib().orc_parse_pattern(text.encode(), int(labeled), C.byref(p))
    if rc:
        raise ValueError(f"oracle: bad pattern ({rc})")
    rc = lib().orc_plan_build(C.byref(p))
    if rc:
        raise ValueError(f"oracle: plan failed ({rc})")
    return p


def _ptrs(rowptr, colidx, labels):
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(colidx, dtype=np.int32)
    lb = None if labels is None else np.ascontiguousarray(labels, dtype=np.int32)
    return rp, ci, lb


def count(rowptr, colidx, labels, p: OrcPlan, edge_lo: int = 0, edge_hi: int = -1, threads: int = 1,
          reuse: bool = True) -> dict:
    """DFS count of the reference semantics.  Returns raw (pre-multiplicity),
    count (raw x |Aut| when unlabeled), alg_bytes and extensions."""
    rp, ci, lb = _ptrs(rowptr, colidx, labels)
    out = (C.c_uint64 * 9)()
    n = rp.shape[0] - 1
    rc = lib().orc_count(n, ci.shape[0], rp.ctypes.data_as(C.POINTER(C.c_int64)),
                         ci.ctypes.data_as(C.POINTER(C.c_int32)),
                         None if lb is None else lb.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(p), edge_lo,
                         edge_hi, threads, int(reuse), out)
    if rc:
        raise ValueError(f"oracle count failed ({rc})")
    raw = int(out[0])
    return {
        "raw": raw,
        "count": raw if p.labeled else raw * int(p.mult),
        "alg_bytes": int(out[1]),
        "extensions": int(sum(out[2:9])),
        "nodes": [int(x) for x in out[2:9]],
    }


def enumerate_embeddings(rowptr, colidx, labels, p: OrcPlan, cap: int = -1) -> tuple[np.ndarray, int]:
    """Every embedding the counting DFS reaches, as rows (M(0), .., M(k-1)) in
    plan order (at most cap rows; -1: all), and the raw count."""
    rp, ci, lb = _ptrs(rowptr, colidx, labels)
    n = rp.shape[0] - 1
    k = int(p.k)
    total = C.c_uint64()
    if cap < 0:
        rc = lib().orc_enumerate(n, ci.shape[0], rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                 ci.ctypes.data_as(C.POINTER(C.c_int32)),
                                 None if lb is None else lb.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(p), None,
                                 0, C.byref(total))
        if rc:
            raise ValueError(f"oracle enumerate failed ({rc})")
        cap = int(total.value)
    out = np.zeros((max(cap, 1), k), dtype=np.int32)
    rc = lib().orc_enumerate(n, ci.shape[0], rp.ctypes.data_as(C.POINTER(C.c_int64)),
                             ci.ctypes.data_as(C.POINTER(C.c_int32)),
                             None if lb is None else lb.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(p),
                             out.ctypes.data_as(C.POINTER(C.c_int32)), cap, C.byref(total))
    if rc:
        raise ValueError(f"oracle enumerate failed ({rc})")
    return out[: min(cap, int(total.value))], int(total.value)


def count_ranges(rowptr, colidx, labels, p: OrcPlan, los, his, threads: int = 1, reuse: bool = True) -> dict:
    """As count(), over the union of edge windows [los[i], his[i])."""
    rp, ci, lb = _ptrs(rowptr, colidx, labels)
    lo = np.ascontiguousarray(los, dtype=np.int64)
    hi = np.ascontiguousarray(his, dtype=np.int64)
    out = (C.c_uint64 * 9)()
    rc = lib().orc_count_ranges(rp.shape[0] - 1, ci.shape[0], rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                ci.ctypes.data_as(C.POINTER(C.c_int32)),
                                None if lb is None else lb.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(p),
                                lo.ctypes.data_as(C.POINTER(C.c_int64)), hi.ctypes.data_as(C.POINTER(C.c_int64)),
                                lo.shape[0], threads, int(reuse), out)
    if rc:
        raise ValueError(f"oracle count failed ({rc})")
    raw = int(out[0])
    return {"raw": raw, "count": raw if p.labeled else raw * int(p.mult), "alg_bytes": int(out[1]),
            "extensions": int(sum(out[2:9]))}


class Prepared:
    """A graph prepared once for many oracle counts (keeps its arrays alive)."""

    def __init__(self, rowptr, colidx, labels=None):
        self.rp, self.ci, self.lb = _ptrs(rowptr, colidx, labels)
        self.h = lib().orc_prepare(self.rp.shape[0] - 1, self.ci.shape[0],
                                   self.rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                   self.ci.ctypes.data_as(C.POINTER(C.c_int32)),
                                   None if self.lb is None else self.lb.ctypes.data_as(C.POINTER(C.c_int32)))

    def count_ranges(self, p: OrcPlan, los, his, threads: int = 1, reuse: bool = True) -> dict:
        lo = np.ascontiguousarray(los, dtype=np.int64)
        hi = np.ascontiguousarray(his, dtype=np.int64)
        out = (C.c_uint64 * 9)()
        rc = lib().orc_count_prepared(self.h, C.byref(p), lo.ctypes.data_as(C.POINTER(C.c_int64)),
                                      hi.ctypes.data_as(C.POINTER(C.c_int64)), lo.shape[0], threads, int(reuse), out)
        if rc:
            raise ValueError(f"oracle count failed ({rc})")
        raw = int(out[0])
        return {"raw": raw, "count": raw if p.labeled else raw * int(p.mult), "alg_bytes": int(out[1]),
                "extensions": int(sum(out[2:9]))}

    def close(self):
        if self.h:
            lib().orc_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bruteforce(rowptr, colidx, labels, p: OrcPlan) -> int:
    """Number of (label-preserving) injective edge-preserving maps P -> G."""
    rp, ci, lb = _ptrs(rowptr, colidx, labels)
    return int(lib().orc_bruteforce_count(rp.shape[0] - 1, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                          ci.ctypes.data_as(C.POINTER(C.c_int32)),
                                          None if lb is None else lb.ctypes.data_as(C.POINTER(C.c_int32)),
                                          C.byref(p)))


def bliss_group(k: int, adj) -> list[tuple[int, ...]] | None:
    """Aut(P) as bliss 0.73 (built from the reference sources) + closure
    returns it; None when oracle/_ref was not built."""
    if not REF.exists():
        return None
    L = C.CDLL(str(REF))
    L.bliss_ref_aut.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
    flat = (C.c_int * (k * k))(*[int(adj[i][j]) for i in range(k) for j in range(k)])
    cap = 5040
    out = (C.c_int * (cap * k))()
    n = L.bliss_ref_aut(k, flat, out, cap)
    if n < 0:
        raise RuntimeError("bliss group too large")
    return [tuple(out[i * k:(i + 1) * k]) for i in range(n)]
