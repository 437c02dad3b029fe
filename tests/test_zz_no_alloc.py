"""tdfs_count_noalloc (the allocation-free variant of tdfs_count, buffers in a
static device array) against tdfs_count, through the C ABI, bit-exact: two
calls in a row, the second on a smaller graph; statistics and edge ranges
too; a plan whose single workgroup's stacks exceed the array is refused."""
import ctypes as C
import os
import time

import numpy as np
import pytest

import tdfs_amd as T
from tdfs_amd import _lib

pytestmark = pytest.mark.gpu
PATTERNS = {
    "triangle": ("v 0 1\nv 1 1\nv 2 1\ne 0 1 1\ne 0 2 1\ne 1 2 1\n", 0),
    "diamond": ("v 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 1 2 1\ne 1 3 1\n", 0),
    "4-clique labeled": ("v 0 0\nv 1 1\nv 2 0\nv 3 1\ne 0 1 1\ne 0 2 1\ne 0 3 1\ne 1 2 1\ne 1 3 1\ne 2 3 1\n", 1),
    "house labeled": ("v 0 0\nv 1 1\nv 2 2\nv 3 0\nv 4 1\ne 0 1 1\ne 1 2 1\ne 2 3 1\ne 3 0 1\ne 2 4 1\ne 3 4 1\n", 1),
    "6-vertex": ("v 0 1\nv 1 1\nv 2 1\nv 3 1\nv 4 1\nv 5 1\ne 0 1 1\ne 1 2 1\ne 0 2 1\ne 2 3 1\ne 1 3 1\n"
                 "e 3 4 1\ne 2 4 1\ne 4 5 1\ne 3 5 1\n", 0),
}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    log = os.environ.get("TDFS_CASE_LOG")
    if log:
        with open(log, "a") as f:
            f.write(" ".join(str(x) for x in a) + "\n")


def noalloc(dev, dg, plan, lo=0, hi=-1):
    """tdfs_count_noalloc through the C ABI: (status, result)."""
    r = _lib.Result()
    rc = _lib.lib().tdfs_count_noalloc(dev.handle, dg.handle, C.byref(plan), lo, hi, C.byref(r))
    return rc, r


def test_noalloc_equals_count():
    dev = T.Device(0)
    big = T.CSRGraph.chung_lu(30000, 150000, 2.1, 1500, 17).with_labels(3, 17)
    small = T.CSRGraph.chung_lu(3000, 12000, 2.1, 200, 18).with_labels(3, 18)
    for gname, G in (("big", big), ("small", small)):  # the second call's input is smaller
        dg = dev.upload(G)
        for name, (text, lab) in PATTERNS.items():
            p = T.plan_pattern(text, lab)
            for stats in (False, True):
                dev.set_stats(stats)
                say("noalloc", gname, name, "stats" if stats else "timed")
                want = dev.count(dg, p)
                rc, got = noalloc(dev, dg, p)
                assert rc == 0, _lib.lib().tdfs_last_error()
                assert got.raw == want.raw and got.count == want.count, (gname, name, stats)
                assert got.flags == want.flags == 0
                if stats:
                    assert got.alg_bytes == want.alg_bytes and got.extensions == want.extensions
            lo, hi = G.nnz // 3, G.nnz // 3 + 777
            assert noalloc(dev, dg, p, lo, hi)[1].raw == dev.count(dg, p, lo, hi).raw
        dev.set_stats(False)
        dg.free()
    dev.close()


def test_noalloc_refuses_stacks_beyond_the_array():
    """A star with 2.2M leaves: a 7-vertex plan needs more DFS stack for one
    workgroup than the array holds (refused, TDFS_E_BUDGET); a triangle needs
    none and is counted (0)."""
    dev = T.Device(0)
    leaves = 2_200_000
    n = leaves + 1
    rowptr = np.zeros(n + 1, dtype=np.int64)
    rowptr[1] = leaves
    rowptr[2:] = leaves + np.arange(1, leaves + 1)
    colidx = np.concatenate([np.arange(1, n, dtype=np.int32), np.zeros(leaves, dtype=np.int32)])
    G = T.CSRGraph(rowptr, colidx)
    t = time.time()
    dg = dev.upload(G)
    say("star uploaded", round(time.time() - t, 1), "s")
    path7 = "".join(f"v {i} 1\n" for i in range(7)) + "".join(f"e {i} {i + 1} 1\n" for i in range(6))
    rc, _ = noalloc(dev, dg, T.plan_pattern(path7, 0))
    assert rc == -7
    assert b"stacks" in _lib.lib().tdfs_last_error()
    rc, r = noalloc(dev, dg, T.plan_pattern(PATTERNS["triangle"][0], 0))
    assert rc == 0 and r.raw == 0
    dg.free()
    dev.close()
