"""Blocks fetched ahead for the slots (claims under load only).

Under load the group whose claim crosses slice 2048 of a slot's block fetches
the slot's next block into the slot's ahead word, and the replacement that
finds the block spent publishes it. Single-lane and deterministic group
replays never fetch ahead: their snapshots, ahead words included, equal the
oracle's (tests/test_gpu_parity.py, tests/test_gpu_group.py). Here: a mixed
run fetches and uses blocks ahead, the blocks held in ahead words are fresh
blocks of the slot's class (no slice claimed), the metadata is quiescent
after the drain, and a pool filled to exhaustion at one size spends the
blocks fetched ahead before slots go empty.
"""
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from gallatin_amd import snapshot as snap
from tests import invariants

pytestmark = pytest.mark.gpu

MiB = 1 << 20
GiB = 1 << 30


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def usable_slots(c, pool_bytes, block_bytes):
    """slots of class c a claim under load may use: 1/16 of the pool in
    blocks, at least the pinned slots, at most 64"""
    pinned = max(4, 128 >> c)
    return min(max(pool_bytes // 16 // block_bytes, pinned), max(pinned, 64))


def ahead_blocks(image):
    s = snap.parse(image)
    a = s["ahead"].astype(np.int64)
    return a[(a > 0) & (a != 0xFFFFFFFF)] - 1, s


def test_mixed_rounds_fetch_and_use_blocks_ahead():
    g = ga.Gallatin.generate_on_device(2 * GiB, 41)
    try:
        before = g.stats()
        g.mixed_reset(1 << 20, 2, 41)
        for r in range(2):
            res = g.mixed_round(r)
            say(f"[ahead] round {r}: {res}")
            assert res["tag_errors"] == 0 and res["alloc_errors"] == 0
        st = g.stats()
        fetched, taken = st["ahead"] - before["ahead"], st["ahead_taken"] - before["ahead_taken"]
        say(f"[ahead] fetched {fetched} taken {taken}")
        assert fetched > 0 and taken > 0
        # a block waiting in an ahead word: a fresh block of its slot's class
        blocks, s = ahead_blocks(g.snapshot())
        assert len(blocks) > 0
        count = s["malloc_counter"].astype(np.int64) & ((1 << 20) - 1)
        assert np.all(count[blocks] == 0) and np.all(s["free_counter"][blocks] == 0)
        g.mixed_drain()
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


def test_exhaustion_spends_blocks_fetched_ahead():
    # 1 GiB at 4 KiB: every block is a whole 16 MiB segment, so blocks held
    # in ahead words are a large share of the pool. Requests for the whole
    # pool run it dry; what is not granted then is the segments the other
    # classes hold (their pinned blocks from boot) and at most one partly
    # claimed block per usable slot of the class. A block left waiting in an
    # ahead word would add to it, and would show in the snapshot.
    pool, size = 1 * GiB, 4096
    g = ga.Gallatin.generate_on_device(pool, 43)
    try:
        n = pool // size
        with ga.Batch(g, n) as b:
            b.set_sizes(np.full(n, size, dtype=np.uint64), seed=43)
            m = b.malloc(width=64, run=1)
            say(f"[ahead] exhaustion: {m}")
            assert m["alloc_errors"] == 0 and m["misses"] > 0
            assert g.fill()["free_segments"] == 0
            blocks, s = ahead_blocks(g.snapshot())
            assert len(blocks) == 0, blocks
            other = int(np.count_nonzero(s["chunk_ids"].astype(np.int64) != 8))
            bound = (other + usable_slots(8, pool, size * 4096)) * 4096
            say(f"[ahead] not granted {n - m['mallocs']}, segments of other classes {other}, bound {bound}")
            assert n - m["mallocs"] <= bound
            b.free()
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()
