"""Pins the whole-allocator oracle: first-principles known answers and invariants.

Known answers are derived from the documented rules, not from the oracle:
the first segment attached is the lowest free one, fresh blocks are granted
in order, a block's slices are handed out in order, recycled blocks come back
in FIFO order, a segment run ends where the top-most qualifying word's low
run of ones ends, and the hashed segment search misses a class's only
segment 0 unless a hashed start lands on it.
"""
import random
import struct

import numpy as np
import pytest

from oracle import oracle
from tests import invariants
from tests.golden.make_murmur_kat import murmur64a as py_murmur

NULL = (1 << 64) - 1
MiB = 1 << 20


def hashed_starts(tid, seed, universe):
    key = struct.pack("<Q", tid)
    h1 = py_murmur(key, seed)
    h2 = py_murmur(key, h1)
    return [((h1 + a * h2) % (1 << 64)) % universe for a in range(15)]


def test_first_slices_come_from_segment_zero_in_order():
    o = oracle.Oracle(1 << 30, 111)
    assert [o.malloc(16, tid=5, slot=0) for _ in range(10)] == [16 * i for i in range(10)]
    assert o.faults() == 0


def test_second_class0_slot_hash_rule():
    # class 0 slot 1 is booted while class 0 owns only segment 0: it stays on
    # segment 0 only if a hashed start (thread id 0, seed + 0) is exactly 0
    for seed in (111, 1, 2, 3, 4, 5, 6, 7):
        o = oracle.Oracle(1 << 30, seed)
        off = o.malloc(16, tid=0, slot=1)
        hits_zero = 0 in hashed_starts(0, seed, 64)
        expect = 1 * 4096 * 16 if hits_zero else 16 * MiB
        assert off == expect, (seed, hits_zero, off)


def test_block_exhaustion_moves_to_next_fresh_block():
    o = oracle.Oracle(1 << 30, 111)
    offs = [o.malloc(16, tid=0, slot=0) for _ in range(4097)]
    assert offs[:4096] == [16 * i for i in range(4096)]
    # the slot's block (segment 0, block 0) is spent; its replacement is the
    # next fresh block of a class-0 segment, so the offset is block aligned
    assert offs[4096] % (16 * 4096) == 0 and offs[4096] != 0
    assert o.faults() == 0


def test_recycled_blocks_return_in_fifo_order():
    # 1 MiB segments, 16..256 B classes. Whole-block requests of 256 KiB take
    # blocks of class 2 (64 B slices, 4 blocks per segment) straight from the
    # class tree, so the block a request gets is observable.
    from gallatin_amd import snapshot as snap
    from tests.test_oracle_veb import expected_find_random, bits_of_words

    seed = 3
    o = oracle.Oracle(64 * MiB, seed, segment=MiB, min_size=16, max_size=256)
    blk = 256 * 1024
    first = [o.malloc(blk, tid=0, slot=0) for _ in range(4)]
    seg = first[0] // MiB
    # fresh blocks of one segment are granted in order
    assert first == [seg * MiB + k * blk for k in range(4)]
    # return three of them in a chosen order; one stays out so the segment
    # keeps its class
    order = [first[2], first[0], first[3]]
    for off in order:
        assert o.free(off) == 0
    # pick thread ids whose hashed search lands on `seg` among the class's
    # listed segments, then the recycled blocks must come back in FIFO order
    listed = set(bits_of_words(snap.parse(o.snapshot())["sub_trees"][2][-1]))
    assert seg in listed
    tids = [t for t in range(5000) if expected_find_random(listed, 64, seed + 2, t) == seg]
    got = [o.malloc(blk, tid=tids[0], slot=0) for _ in range(3)]
    assert got == order
    assert o.faults() == 0


def test_segment_run_ends_at_top_word_run():
    o = oracle.Oracle(4 << 30, 9)  # 256 segments, 4 words; boot takes low segments
    off = o.malloc(17 * MiB, tid=0, slot=0)
    assert off == 254 * 16 * MiB
    off2 = o.malloc(40 * MiB, tid=0, slot=0)  # 3 segments, below the first run
    assert off2 == 251 * 16 * MiB
    assert o.free(off) == 0
    # segments 254-255 are free again but sit above the held 251-253, outside
    # the word's low run of ones (bits 0..58), so the next run is 249-250
    off3 = o.malloc(32 * MiB, tid=0, slot=0)
    assert off3 == 249 * 16 * MiB
    assert o.free(off2) == 0 and o.free(off3) == 0
    off4 = o.malloc(17 * MiB, tid=0, slot=0)  # whole word free: top of it again
    assert off4 == 254 * 16 * MiB
    assert o.faults() == 0


def test_segment_requests_fail_on_one_word_tree():
    o = oracle.Oracle(1 << 30, 9)  # 64 segments: one word, never served
    assert o.malloc(17 * MiB) == NULL


def test_block_path_offsets_are_block_aligned():
    o = oracle.Oracle(1 << 30, 4)
    for size in (33 * 1024, 64 * 1024, 1 * MiB, 16 * MiB):
        off = o.malloc(size, tid=3, slot=0)
        _, tree, _ = oracle.route(16 * MiB, 16, 4096, size)
        assert off != NULL and off % ((16 << tree) * 4096) == 0


def random_sequence(rng, n, sizes, free_p=0.45):
    live, ops, results_idx = [], [], []
    for i in range(n):
        if live and rng.random() < free_p:
            j = live.pop(rng.randrange(len(live)))
            ops.append((1, j, 0, 0))
        else:
            ops.append((0, rng.choice(sizes), rng.randrange(1 << 20), rng.randrange(256)))
            live.append(len(ops) - 1)
    return ops


def replay_and_check(o, ops, pool, seg, mn, mx):
    # replay op by op so that failed mallocs are not freed
    res = []
    live = {}
    for i, (kind, arg, tid, slot) in enumerate(ops):
        if kind == 0:
            off = o.malloc(arg, tid, slot)
            res.append(off)
            if off != NULL:
                live[i] = (off, arg)
        else:
            if arg in live:
                off, _ = live.pop(arg)
                assert o.free(off) == 0
            res.append(0)
        if i % 97 == 0 and live:
            offs = np.array([v[0] for v in live.values()], dtype=np.uint64)
            sizes = np.array([v[1] for v in live.values()], dtype=np.uint64)
            invariants.check_live(offs, sizes, pool, seg, mn, mx, oracle.route)
    return res, live


@pytest.mark.parametrize(
    "pool,seg,mn,mx,sizes",
    [
        (1 << 30, 16 * MiB, 16, 4096, [16, 24, 32, 100, 512, 4096]),
        (1 << 30, 16 * MiB, 16, 4096, [16, 5000, 8192, 12000, 16384, 32768]),  # big slices
        (4 << 30, 16 * MiB, 16, 4096, [16, 64 * 1024, 1 * MiB, 16 * MiB, 17 * MiB, 40 * MiB]),
        (64 * MiB, MiB, 16, 256, [16, 32, 256, 1024, 8192, 65536, 2 * MiB]),  # exhaustion
    ],
)
def test_random_sequences_hold_invariants(pool, seg, mn, mx, sizes):
    rng = random.Random(pool ^ len(sizes))
    o = oracle.Oracle(pool, 77, segment=seg, min_size=mn, max_size=mx)
    ops = random_sequence(rng, 6000, sizes)
    res, live = replay_and_check(o, ops, pool, seg, mn, mx)
    for off, _ in live.values():
        assert o.free(off) == 0
    invariants.check_quiescent(o.snapshot(), seg, mn)
    assert o.faults() == 0


def test_out_of_memory_then_recovery():
    # fill a small pool completely, free everything, allocate again
    pool, seg = 64 * MiB, MiB
    o = oracle.Oracle(pool, 5, segment=seg, min_size=16, max_size=256)
    got = []
    for i in range(400000):
        off = o.malloc(256, tid=i, slot=i % 64)
        if off == NULL:
            break
        got.append(off)
    assert off == NULL and len(got) > 1000  # ran out
    for off in got:
        assert o.free(off) == 0
    invariants.check_quiescent(o.snapshot(), seg, 16)
    again = [o.malloc(256, tid=i, slot=i % 64) for i in range(len(got) // 2)]
    assert NULL not in again  # the class came back after the frees
    assert o.faults() == 0


def test_free_rejects_foreign_offsets():
    o = oracle.Oracle(1 << 30, 1)
    assert o.free(1 << 40) != 0      # outside the pool
    assert o.free(63 * 16 * MiB) != 0  # a segment nobody owns
