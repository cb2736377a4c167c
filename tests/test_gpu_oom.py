"""Out of memory under concurrency, on the GPU.

The hashed block search never reaches segment 0 from a start above it
(reference veb.cuh:960-975, kept for parity), and segment 0 belongs to class
0 (16 B slices) from construction on. Once the pool is exhausted and class
0's other segments are used up, a class-0 block request fails almost always
(only a hashed start of exactly 0 finds segment 0), while the class still
looks occupied. A group that finds every pinned slot of its class empty must
then report out of memory unless its own request could succeed: asking for a
refill that can only fail made each wave queue behind the empty slots in
turn (one failed refill per wave; on the mixed 8 GiB workload a kernel took
a hundred times its normal duration).
"""
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from gallatin_amd import snapshot as snap
from oracle import oracle

pytestmark = pytest.mark.gpu

NULL = (1 << 64) - 1
MiB = 1 << 20
GiB = 1 << 30


def items(tree):
    """Items of a parsed vEB tree (its bottom layer's set bits)."""
    out = []
    for w, x in enumerate(tree[-1]):
        x = int(x)
        while x:
            b = (x & -x).bit_length() - 1
            out.append(w * 64 + b)
            x &= x - 1
    return out


def exhausted_pool(pool, seed=42):
    """A Gallatin<16 MiB, 16, 4096> whose free segments are all taken by
    single-lane 16 MiB mallocs (one class-8 block = one segment each), checked
    against the oracle; returns the allocator and the oracle's class-0 list."""
    o = oracle.Oracle(pool, seed)
    ops = [(0, 16 * MiB, 0, 0)] * (pool // (16 * MiB))
    want = o.replay(ops)
    assert want[-1] == NULL  # boot holds some segments: the last malloc fails
    s = snap.parse(o.snapshot())
    assert items(s["segment_tree"]) == []
    g = ga.Gallatin.generate_on_device(pool, seed)
    got = g.replay(ops)
    assert np.array_equal(got, want)
    assert np.array_equal(g.snapshot(), o.snapshot())
    return g, items(s["sub_trees"][0])


def test_class_with_unreachable_segment_reports_oom_without_refill_queue():
    print("[oom] 8 GiB pool exhausted by 16 MiB mallocs; 2^21 mallocs of 16 B in one kernel", flush=True)
    sys.stdout.flush()
    g, class0 = exhausted_pool(8 * GiB)
    try:
        # segment 0 (hashed starts rarely reach it) and segment 1 (reached by
        # the search's final successor step) are class 0's
        assert class0 == [0, 1]
        n = 1 << 21
        before = g.stats()
        out, r = g.churn(np.zeros(n, np.uint32), np.full(n, 16, np.uint64), seed=7)
        after = g.stats()
        refills = after["refill"] - before["refill"]
        ok = out[out != NULL]
        print(f"[oom] {r} refills {refills} granted {len(ok)}", flush=True)
        assert r["mallocs"] == len(ok) and r["mallocs"] + r["misses"] == n
        assert r["alloc_errors"] == 0 and g.errors() == (0, 0)
        assert len(np.unique(ok)) == len(ok)
        assert (ok % 16 == 0).all() and (ok < 8 * GiB).all()
        # the pinned blocks and segment 1 are used up: misses follow
        assert r["misses"] > 0
        # Refills are asked for only by groups whose own request can succeed
        # (a hashed start of 0: about 3% of them), not by every wave that
        # meets an empty slot (2^15 waves; the per-wave queue did about 4000)
        assert refills <= 1000, refills
    finally:
        g.free_on_device()
