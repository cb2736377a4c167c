"""Pins the oracle's model of malloc under load (GAL_OP_MALLOC_SPREAD).

When more than one lane of a wave claims, the GPU's first claim pass of a
malloc call starts each class group at slot hint % usable of its class
instead of the reference's pinned-slot walk (gallatin.cuh:865-1078 walks
shared_block_storage.cuh:119-145 from hint % pinned). `usable` counts the
class's pinned slots and overflow slots up to 64, as many as hold 1/16 of
the pool in blocks, at least the pinned ones. A slot holding a block is
taken; an empty one is refilled unless the class is out of memory. What
fixes the results, independently of the oracle's team code:
  * one lane never takes the look: a one-lane spread call site is the
    single-lane path;
  * a hint below the class's pinned count lands where the walk starts: on a
    slot holding a block the spread call site equals the plain one;
  * the usable count is a closed form of the pool and the class (below):
    hint usable - 1 lands on that slot, hint usable wraps to slot 0;
  * a look at an empty overflow slot fills it with a fresh block of the
    class that no other slot holds (the slot image shows it there, no slice
    taken); the group itself claims in its next pass, which takes no look:
    from the pinned walk, as a plain call site would;
  * the next spread call site with that hint takes the overflow block's
    first slices in lane order; a plain call site with the same hint walks
    the pinned slots and never sees it;
  * random streams of spread and plain call sites keep the allocator
    invariants and replay to the same results.
"""
import random

import numpy as np
import pytest

from gallatin_amd import snapshot as snap
from oracle import oracle
from tests import group_streams, invariants

NULL = (1 << 64) - 1
MiB = 1 << 20
GiB = 1 << 30


def lanes_of(mask):
    return [lane for lane in range(64) if (mask >> lane) & 1]


def pinned(c):
    return max(4, 128 >> c)


def region(c):
    """slots of class c in the slot image: pinned ones, then overflow up to 64"""
    return max(pinned(c), 64)


def region_start(c):
    return sum(region(k) for k in range(c))


def usable(c, pool, seg=16 * MiB, mn=16):
    block = (mn << c) * 4096
    return min(max(pool // 16 // block, pinned(c)), region(c))


def class_slots(image, c):
    s = snap.parse(image)["slots"]
    return s[region_start(c):region_start(c) + region(c)].copy()


def test_usable_closed_form_examples():
    # 1 GiB pool, 16 MiB segments: the 4 KiB class (16 MiB blocks) has 4,
    # the 512 B class (2 MiB blocks) 32, the 64 B class its 64, class 0 its 128
    assert [usable(c, GiB) for c in (0, 1, 2, 5, 8)] == [128, 64, 64, 32, 4]
    assert [usable(c, 8 * GiB) for c in (5, 6, 7, 8)] == [64, 64, 64, 32]


def test_one_lane_spread_is_the_single_lane_path():
    rng = random.Random(3)
    a, b = oracle.Oracle(GiB, 5), oracle.Oracle(GiB, 5)
    for _ in range(2000):
        lane, size = rng.randrange(64), rng.choice([16, 64, 512, 1024, 4096, 8192])
        slot, tid = rng.randrange(256), rng.randrange(1 << 20)
        sizes, tids = [size] * 64, [tid] * 64
        x = a.group_malloc(1 << lane, sizes, tids, slot, spread=True)[lane]
        y = b.group_malloc(1 << lane, sizes, tids, slot, spread=False)[lane]
        assert x == y
    assert np.array_equal(a.snapshot(), b.snapshot())


@pytest.mark.parametrize("c", [2, 4, 5, 6, 7])
def test_hint_below_pinned_equals_the_walk(c):
    size = 16 << c
    for hint in (0, 1, pinned(c) - 1):
        a, b = oracle.Oracle(GiB, 9), oracle.Oracle(GiB, 9)
        lanes = 0b1011_0110
        x = a.group_malloc(lanes, [size] * 64, list(range(64)), hint, spread=True)
        y = b.group_malloc(lanes, [size] * 64, list(range(64)), hint, spread=False)
        assert x == y
        assert np.array_equal(a.snapshot(), b.snapshot())


@pytest.mark.parametrize("pool,c", [(GiB, 5), (GiB, 6), (GiB, 7), (2 * GiB, 8), (8 * GiB, 8), (8 * GiB, 3)])
def test_look_fills_the_overflow_slot_hint_mod_usable(pool, c):
    u = usable(c, pool)
    assert u > pinned(c)
    size = 16 << c
    block_bytes = size * 4096
    lanes = (1 << 2) | (1 << 9) | (1 << 10) | (1 << 40)
    for hint in (u - 1, pinned(c) + (u - pinned(c)) // 2, 3 * u + pinned(c)):
        si = hint % u
        assert si >= pinned(c)
        o, twin = oracle.Oracle(pool, 17), oracle.Oracle(pool, 17)
        before = class_slots(o.snapshot(), c)
        assert before[si] == 0 and all(before[:pinned(c)] != 0)
        out = o.group_malloc(lanes, [size] * 64, [7] * 64, hint, spread=True)
        image = o.snapshot()
        after = class_slots(image, c)
        # only the looked-at slot changed, and it holds a block no other slot held
        changed = np.nonzero(after != before)[0].tolist()
        assert changed == [si]
        block = int(after[si]) - 1
        assert int(after[si]) not in before.tolist()
        s = snap.parse(image)
        bps = s["blocks_per_segment"]
        assert int(s["chunk_ids"][block // bps]) == c  # its segment serves the class
        assert int(s["malloc_counter"][block]) == c << 20  # no slice taken from it
        # the group claimed in its next pass, by the pinned walk
        walk = twin.group_malloc(lanes, [size] * 64, [7] * 64, hint, spread=False)
        assert [out[lane] for lane in lanes_of(lanes)] == [walk[lane] for lane in lanes_of(lanes)]
        # the next spread call site with that hint takes the overflow block's first slices
        start = (block // bps) * (16 * MiB) + (block % bps) * block_bytes
        more = o.group_malloc(0b1101, [size] * 64, [7] * 64, hint, spread=True)
        assert [more[0], more[2], more[3]] == [start, start + size, start + 2 * size]
        # a plain call site with the same hint walks the pinned slots instead
        plain = o.group_malloc(0b1, [size] * 64, [7] * 64, hint, spread=False)[0]
        assert not start <= plain < start + block_bytes


@pytest.mark.parametrize("c", [5, 6, 7])
def test_hint_usable_wraps_to_slot_zero(c):
    u = usable(c, GiB)
    size = 16 << c
    a, b = oracle.Oracle(GiB, 4), oracle.Oracle(GiB, 4)
    lanes = 0xFF00
    x = a.group_malloc(lanes, [size] * 64, [1] * 64, u, spread=True)
    y = b.group_malloc(lanes, [size] * 64, [1] * 64, 0, spread=False)
    assert x == y
    assert np.array_equal(a.snapshot(), b.snapshot())


def test_out_of_memory_class_fails_at_once():
    # 256 MiB pool: exhaust the 4 KiB class one lane at a time until every
    # pinned slot holds 0 and no segment is free; a spread call site of
    # that class then fails for all its lanes and changes nothing
    o = oracle.Oracle(256 * MiB, 2)
    misses = 0
    for i in range(200000):
        if o.malloc(4096, i, i % 7) == NULL:
            misses += 1
            if misses > 64:
                break
    image = o.snapshot()
    assert all(class_slots(image, 8)[:pinned(8)] == 0)
    assert snap.popcount64(snap.parse(image)["segment_tree"][-1]) == 0
    out = o.group_malloc((1 << 64) - 1, [4096] * 64, list(range(64)), 3, spread=True)
    assert all(x == NULL for x in out)
    assert np.array_equal(o.snapshot(), image)


SPREAD_CASES = [
    ("slices", GiB, (16 * MiB, 16, 4096), [16, 32, 48, 64, 256, 1000, 2048, 4096], (), 500),
    ("big_slices", 2 * GiB, (16 * MiB, 16, 4096), [64, 512, 4096, 8192, 16384, 32768], (), 400),
    ("large", 4 * GiB, (16 * MiB, 16, 4096), [16, 256, 4096], (40000, MiB, 40 * MiB), 300),
    ("exhaustion", 64 * MiB, (MiB, 16, 256), [16, 32, 256, 1024, 4096], (65536, 2 * MiB), 600),
    ("small_pool_oom", 256 * MiB, (16 * MiB, 16, 4096), [1024, 2048, 4096], (), 600),
]


@pytest.mark.parametrize("name,pool,triple,sizes,large,n", SPREAD_CASES, ids=[c[0] for c in SPREAD_CASES])
def test_random_spread_streams_keep_the_invariants(name, pool, triple, sizes, large, n):
    seg, mn, mx = triple
    o = oracle.Oracle(pool, 6, segment=seg, min_size=mn, max_size=mx)
    ops, results, live = group_streams.build(o, 6, n, sizes, large, seg=seg, mn=mn, mx=mx,
                                             large_p=0.08 if large else 0.0, spread_p=0.7)
    assert o.faults() == 0
    assert any(op[0] == 2 for op in ops) and any(op[0] == 0 for op in ops)
    size_of = {}
    for i, op in enumerate(ops):
        if op[0] in (0, 2):
            for lane in lanes_of(op[2]):
                size_of[i * 64 + lane] = op[3][lane]
    offs = np.array([e[1] for e in live], dtype=np.uint64)
    szs = np.array([size_of[e[0]] for e in live], dtype=np.uint64)
    invariants.check_live(offs, szs, pool, seg, mn, mx, oracle.route)
    again = oracle.Oracle(pool, 6, segment=seg, min_size=mn, max_size=mx).group_replay(ops)
    assert np.array_equal(again, np.array(results, dtype=np.uint64))
    for e in live:
        assert o.free(e[1]) == 0
    invariants.check_quiescent(o.snapshot(), seg, mn)
    assert o.faults() == 0
