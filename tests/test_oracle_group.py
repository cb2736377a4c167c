"""Pins the oracle's multi-lane (one warp call site) model.

The reference claims slices per coalesced team of a warp
(gallatin.cuh:865-1078, block.cuh:245-281) and frees per team of one block
(gallatin.cuh:1774-1827). What fixes the results, independently of the
oracle's team code:
  * a call site of one lane is the single-lane path (already pinned);
  * a team's claim is one bump of the block counter by the team total, each
    lane taking the exclusive prefix of the counts in lane order: when no
    range crosses slice 4095 this equals the lanes' single-lane mallocs made
    one after another in lane order (same offsets, same metadata);
  * teams of different classes use different slots and blocks: a mixed call
    site equals its lanes' single-lane mallocs in lane order when nothing is
    replaced;
  * a range that straddles slice 4095 fails, its lane retries on the
    replacement block, and the slices no lane holds are written off (closed
    form below);
  * a team free adds the team size to the block's free counter once: equal to
    the lanes' single frees;
  * random streams keep the allocator invariants (no overlap, alignment,
    in-pool, quiescent metadata after everything is freed).
"""
import random

import numpy as np
import pytest

from gallatin_amd import snapshot as snap
from oracle import oracle
from tests import group_streams, invariants

NULL = (1 << 64) - 1
MiB = 1 << 20
FULL = (1 << 64) - 1


def lanes_of(mask):
    return [lane for lane in range(64) if (mask >> lane) & 1]


def test_one_lane_call_sites_are_the_single_lane_path():
    rng = random.Random(5)
    sizes = [16, 32, 100, 4096, 8192, 20000, 40000, MiB, 17 * MiB]
    a, b = oracle.Oracle(1 << 30, 9), oracle.Oracle(1 << 30, 9)
    ops, single = [], []
    live = []
    for i in range(3000):
        lane = rng.randrange(64)
        if live and rng.random() < 0.4:
            j, single_j = live.pop(rng.randrange(len(live)))
            args = [0] * 64
            args[lane] = j
            ops.append((1, 0, 1 << lane, args, [0] * 64))
            single.append((1, single_j, 0, 0))
        else:
            size, tid, slot = rng.choice(sizes), rng.randrange(1 << 20), rng.randrange(256)
            args, tids = [0] * 64, [0] * 64
            args[lane], tids[lane] = size, tid
            ops.append((0, slot, 1 << lane, args, tids))
            single.append((0, size, tid, slot))
            if b.malloc(size, tid, slot) != NULL:  # b only tells which mallocs succeed
                live.append(((len(ops) - 1) * 64 + lane, len(single) - 1))
    got = a.group_replay(ops)
    expect = oracle.Oracle(1 << 30, 9).replay(single)
    flat = np.array([got[i][lanes_of(op[2])[0]] for i, op in enumerate(ops)], dtype=np.uint64)
    assert np.array_equal(flat, expect)
    ref = oracle.Oracle(1 << 30, 9)
    ref.replay(single)
    assert np.array_equal(a.snapshot(), ref.snapshot())


def test_team_prefix_in_lane_order_on_a_fresh_allocator():
    # class 0's slot 0 holds block 0 of segment 0 from boot, with no slice out
    lanes = (1 << 3) | (1 << 7) | (1 << 8) | (1 << 20) | (1 << 63)
    o = oracle.Oracle(1 << 30, 111)
    out = o.group_malloc(lanes, [16] * 64, [0] * 64, 0)
    assert [out[lane] for lane in lanes_of(lanes)] == [16 * r for r in range(5)]
    # the next single lane continues after the team's total
    assert o.malloc(16, 0, 0) == 16 * 5


def test_multi_slice_counts_take_their_prefix():
    # 16 KiB requests take 4 top-class slices, 8 KiB two, 4 KiB one: each lane
    # starts at the sum of the counts of the lanes below it
    o = oracle.Oracle(1 << 30, 3)
    base = o.malloc(4096, 0, 0)
    sizes = [4096, 16384, 8192, 4096, 16384, 8192] + [4096] * 58
    lanes = 0b111111
    out = o.group_malloc(lanes, sizes, [0] * 64, 0)
    counts = [1, 4, 2, 1, 4, 2]
    prefix = np.concatenate([[0], np.cumsum(counts)[:-1]])
    assert [out[lane] for lane in range(6)] == [base + 4096 * (1 + int(p)) for p in prefix]
    assert o.malloc(4096, 0, 0) == base + 4096 * (1 + sum(counts))


@pytest.mark.parametrize("seed", range(6))
def test_team_equals_lanes_in_order_when_nothing_is_replaced(seed):
    rng = random.Random(seed)
    sizes_pool = [16, 32, 48, 64, 128, 256, 1000, 2048, 4096, 8192, 16384]
    a, b = oracle.Oracle(2 << 30, seed), oracle.Oracle(2 << 30, seed)
    checked = 0
    for step in range(300):
        lanes = group_streams.lane_mask(rng)
        slot = rng.randrange(256)
        one = rng.random() < 0.5
        sizes = [rng.choice(sizes_pool)] * 64 if one else [rng.choice(sizes_pool) for _ in range(64)]
        tids = [rng.randrange(1 << 20) for _ in range(64)]
        seq = [b.malloc(sizes[lane], tids[lane], slot) for lane in lanes_of(lanes)]
        got = a.group_malloc(lanes, sizes, tids, slot)
        got = [got[lane] for lane in lanes_of(lanes)]
        stats_same = np.array_equal(a.snapshot(), b.snapshot())
        if got == seq and stats_same:
            checked += 1
            continue
        # they may differ only when the call site crossed slice 4095 of a
        # block (then the team retries on the replacement, the sequential
        # lanes do not write anything off): resynchronise and go on
        a_snap = snap.parse(a.snapshot())
        spent = np.nonzero((a_snap["malloc_counter"] & 0xFFFFF) >= 4096)[0]
        assert len(spent) > 0, (step, got, seq)
        a, b = oracle.Oracle(2 << 30, seed + 100 + step), oracle.Oracle(2 << 30, seed + 100 + step)
    assert checked > 200


def test_straddling_lane_retries_on_the_replacement_and_the_rest_is_written_off():
    # class 8 (4 KiB slices): fill slot 0's block to slice 4093 one lane at a time
    o = oracle.Oracle(1 << 30, 21)
    first = [o.malloc(4096, 0, 0) for _ in range(4093)]
    block_start = first[0]
    assert first == [block_start + 4096 * k for k in range(4093)]
    block = block_start // (16 * MiB)  # one 4 KiB-class block per 16 MiB segment
    # three lanes of 8 KiB (two slices each): lane 0 takes 4093-4094, lane 1
    # would take 4095-4096 (straddles), lane 2 4097-4098 (outside)
    out = o.group_malloc(0b111, [8192] * 64, [0] * 64, 0)
    assert out[0] == block_start + 4096 * 4093
    s = snap.parse(o.snapshot())
    blk = int(block) * s["blocks_per_segment"]
    assert int(s["malloc_counter"][blk]) & 0xFFFFF == 4093 + 6
    # written off: in-range slices (3) minus the lanes that hold one (1) plus
    # lane 0's second slice ... i.e. 3 - 1 = 2 slices
    assert int(s["free_counter"][blk]) == 2
    # lanes 1 and 2 retried together on the fresh block: prefix 0 and 2
    fresh = out[1]
    assert fresh % (16 * MiB) == 0 and fresh != block_start
    assert out[2] == fresh + 8192


def test_team_free_equals_single_frees():
    rng = random.Random(7)
    a, b = oracle.Oracle(1 << 30, 2), oracle.Oracle(1 << 30, 2)
    offs = []
    for _ in range(6000):
        size, slot = rng.choice([16, 64, 4096, 8192]), rng.randrange(8)
        offs.append(a.malloc(size, 0, slot))
        assert b.malloc(size, 0, slot) == offs[-1]
    order = list(range(len(offs)))
    rng.shuffle(order)
    while order:
        k = rng.randint(1, 64)
        batch, order = order[:k], order[k:]
        lanes = rng.sample(range(64), len(batch))
        mask, arr = 0, [0] * 64
        for lane, j in zip(lanes, batch):
            mask |= 1 << lane
            arr[lane] = offs[j]
        assert a.group_free(mask, arr) == 0
        for lane in sorted(lanes):
            assert b.free(arr[lane]) == 0
        assert np.array_equal(a.snapshot(), b.snapshot())
    invariants.check_quiescent(a.snapshot(), 16 * MiB, 16)


CASES = [
    ("slices", 1 << 30, (16 * MiB, 16, 4096), [16, 32, 48, 64, 256, 1000, 2048, 4096], (), 400),
    ("big_slices", 1 << 30, (16 * MiB, 16, 4096), [16, 4096, 8192, 12000, 16384, 32768], (), 400),
    ("large", 4 << 30, (16 * MiB, 16, 4096), [16, 256, 4096, 8192], (40000, MiB, 16 * MiB, 40 * MiB), 300),
    ("exhaustion", 64 * MiB, (MiB, 16, 256), [16, 32, 256, 1024, 4096], (8192, 65536, 2 * MiB), 600),
]


@pytest.mark.parametrize("name,pool,triple,sizes,large,n", CASES, ids=[c[0] for c in CASES])
def test_random_group_streams_keep_the_invariants(name, pool, triple, sizes, large, n):
    seg, mn, mx = triple
    o = oracle.Oracle(pool, 4, segment=seg, min_size=mn, max_size=mx)
    ops, results, live = group_streams.build(o, 4, n, sizes, large, seg=seg, mn=mn, mx=mx,
                                             large_p=0.08 if large else 0.0)
    assert o.faults() == 0
    offs = np.array([e[1] for e in live], dtype=np.uint64)
    size_of = {}
    for i, op in enumerate(ops):
        if op[0] == 0:
            for lane in lanes_of(op[2]):
                size_of[i * 64 + lane] = op[3][lane]
    szs = np.array([size_of[e[0]] for e in live], dtype=np.uint64)
    invariants.check_live(offs, szs, pool, seg, mn, mx, oracle.route)
    # the whole stream replayed on a fresh oracle gives the same results
    again = oracle.Oracle(pool, 4, segment=seg, min_size=mn, max_size=mx).group_replay(ops)
    assert np.array_equal(again, np.array(results, dtype=np.uint64))
    for e in live:
        assert o.free(e[1]) == 0
    invariants.check_quiescent(o.snapshot(), seg, mn)
    assert o.faults() == 0
