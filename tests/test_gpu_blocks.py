"""The block and multi-segment path under concurrent churn (gal_churn), at
the 64 GiB pool of the block/multi-segment workload with fewer requests.

Every round: no tag clash, no internal error, the live set satisfies the
live invariants, and the fragmentation figure is right end to end: one
malloc of exactly largest-run segments succeeds and one of a segment more
fails (pinned against the oracle's gather rule on the segment tree's words
too). After the drain the allocator is quiescent.
"""
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from gallatin_amd import snapshot as snap
from gallatin_amd.workloads import BlockChurn
from oracle import oracle
from tests import invariants

pytestmark = pytest.mark.gpu

MiB = 1 << 20
SEG = 16 * MiB


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def one(g, kind, arg):
    res, r = g.churn([kind], [arg], seed=99)
    return int(res[0]), r


def oracle_gathers(words, n):
    t = oracle.Veb(len(words) * 64)
    t.set_words(t.layers() - 1, words)
    return t.gather_multiple(n) != oracle.NULL


def test_block_churn_rounds_and_fragmentation():
    pool = 64 << 30
    g = ga.Gallatin.generate_on_device(pool, 21)
    try:
        # runs are claimed from the top of the segment tree and class segments
        # from the bottom: the largest run shrinks by about 500 segments a
        # round and gives out (misses) from about round 8 on
        w = BlockChurn(g, requests=4096, rounds=11, seed=5)
        for _ in range(w.rounds):
            rec = w.step()
            say(f"[block-churn] {rec}")
            assert rec["tag_errors"] == 0 and rec["alloc_errors"] == 0, rec
            offs, sizes = w.live_arrays()
            invariants.check_live(offs, sizes, pool, SEG, 16, 4096, oracle.route)
            n = rec["largest_run_segments"]
            words = snap.parse(g.snapshot())["segment_tree"][-1]
            assert ga.gather_limit(words) == n
            assert (n == 0 or oracle_gathers(words, n)) and not oracle_gathers(words, n + 1)
            # end to end: the largest run is exactly what one malloc can get
            if n >= 1:
                off, r = one(g, ga.GAL_OP_MALLOC, n * SEG)
                assert off != ga.GAL_NULL_OFFSET and r["tag_errors"] == 0, (n, r)
                one(g, ga.GAL_OP_FREE, off)
            off, r = one(g, ga.GAL_OP_MALLOC, (n + 1) * SEG)
            assert off == ga.GAL_NULL_OFFSET, (n, off)
        assert sum(h["mallocs"] for h in w.history) > 0 and sum(h["frees"] for h in w.history) > 0
        assert any(h["misses"] > 0 for h in w.history), "the workload should run the pool out"
        r = w.drain()
        say(f"[block-churn] drain {r}")
        assert r is None or r["tag_errors"] == 0
        assert g.errors() == (0, 0), g.error_sites()
        s = invariants.check_quiescent(g.snapshot(), SEG, 16)
        assert g.free_bytes() == snap.popcount64(s["segment_tree"][-1]) * SEG
        assert g.largest_segment_run() > 0
    finally:
        g.free_on_device()


@pytest.mark.parametrize("size", [64 << 10, 3 << 20, 16 << 20, 17 << 20, 100 << 20])
def test_churn_same_size_many_lanes(size):
    """Many lanes of one wave asking for blocks or segment runs at once."""
    pool = 8 << 30
    g = ga.Gallatin.generate_on_device(pool, 22)
    try:
        n = 512
        res, r = g.churn(np.zeros(n, dtype=np.uint32), np.full(n, size, dtype=np.uint64), seed=3)
        ok = res != ga.GAL_NULL_OFFSET
        say(f"[churn-same] size={size} got {int(ok.sum())}/{n} {r}")
        assert r["tag_errors"] == 0 and r["alloc_errors"] == 0
        assert ok.sum() > 0
        invariants.check_live(res[ok], np.full(int(ok.sum()), size, dtype=np.uint64), pool, SEG, 16, 4096,
                              oracle.route)
        _, r2 = g.churn(np.ones(int(ok.sum()), dtype=np.uint32), res[ok], seed=3)
        assert r2["tag_errors"] == 0 and r2["frees"] == int(ok.sum())
        assert g.errors() == (0, 0), g.error_sites()
        invariants.check_quiescent(g.snapshot(), SEG, 16)
    finally:
        g.free_on_device()
