"""The pinned-slot walk done by a group's lanes together
(Gallatin::pinned_slot_group) must stop where the reference's sequential walk
(shared_block_storage.cuh:119-145, restated in
the oracle) stops: same slot, same entry, for every slot count, start and
group size. Both walks are restated here on random slot states, including
all-empty, all-busy and wrapping starts; a single-lane group is the
sequential walk itself, which the GPU replay tests compare with the oracle.
"""
import random

import pytest

BUSY = -1
CAP = 30  # SHARED_BLOCK_COUNTER_CUTOFF


def sequential(slots, hint):
    n = len(slots)
    my = hint % n
    original = my
    entry = slots[my]
    saw = entry == BUSY
    counter = 0
    while (entry == 0 or entry == BUSY) and my != original - 1:
        if saw and counter + 1 >= n:
            break
        my = (my + 1) % n
        entry = slots[my]
        saw |= entry == BUSY
        counter += 1
        if counter >= CAP:
            break
    if saw and (entry == 0 or entry == BUSY):
        entry = BUSY
    return my, entry


def by_group(slots, hint, g):
    """Positions base .. base+g-1 loaded at once; the first stopping position wins."""
    n = len(slots)
    original = hint % n
    last = n - 1 if (original >= 1 and n - 1 < CAP) else CAP
    saw = False
    base = 0
    while True:
        ks = [base + r for r in range(g)]
        es = [slots[(original + k) % n] if k <= last else 0 for k in ks]
        busy = [k <= last and e == BUSY for k, e in zip(ks, es)]
        for r, (k, e) in enumerate(zip(ks, es)):
            if k > last:
                continue
            busy_so_far = saw or any(busy[: r + 1])
            if (e != 0 and e != BUSY) or k == last or (busy_so_far and k + 1 >= n):
                return (original + k) % n, (BUSY if busy_so_far and e in (0, BUSY) else e)
        saw = saw or any(busy)
        base += g


@pytest.mark.parametrize("seed", range(4))
def test_group_walk_equals_sequential_walk(seed):
    rng = random.Random(seed)
    for _ in range(5000):
        n = rng.choice([4, 8, 16, 32, 64, 128])
        p_valid = rng.choice([0.0, 0.02, 0.1, 0.5])
        p_busy = rng.choice([0.0, 0.05, 0.3, 1.0])
        slots = [
            BUSY if rng.random() < p_busy else (rng.randrange(1, 1000) if rng.random() < p_valid else 0)
            for _ in range(n)
        ]
        hint = rng.randrange(1024)
        g = rng.randrange(1, 65)
        assert by_group(slots, hint, g) == sequential(slots, hint), (n, hint, g, slots)


def test_known_walks():
    # start 0 never meets "original - 1": the walk wraps until the cap
    assert sequential([0, 0, 0, 0], 0) == (30 % 4, 0)
    # a busy slot ends the walk after one full cycle, reporting busy
    assert sequential([0, BUSY, 0, 0], 2) == (1, BUSY)
    # the first slot holding a block, in walk order
    assert sequential([0, 0, 7, 9], 0) == (2, 7)
    assert sequential([0, 0, 7, 9], 3) == (3, 9)
    for g in (1, 2, 3, 64):
        assert by_group([0, 0, 0, 0], 0, g) == (30 % 4, 0)
        assert by_group([0, BUSY, 0, 0], 2, g) == (1, BUSY)
        assert by_group([0, 0, 7, 9], 0, g) == (2, 7)
        assert by_group([0, 0, 7, 9], 3, g) == (3, 9)
