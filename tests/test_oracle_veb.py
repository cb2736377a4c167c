"""Pins the oracle's vEB tree against brute-force models on small universes.

A Python set is the model of the item layer. With only insert/remove applied
the summaries stay exact, so successor(x) must be min{y in S : y > x}. The
hashed start of find_random_valid_index is recomputed with an independent
MurmurHash64A. gather_multiple is checked against a declarative statement of
its rule (the highest word w >= 1 whose low run of ones, extended downward
through full words and the high bits of the next word, reaches n; the run
ends at w*64 + trailing_ones(w)), and a failed gather must leave every word
unchanged.
"""
import random
import struct

import numpy as np
import pytest

from oracle import oracle
from tests.golden.make_murmur_kat import murmur64a as py_murmur

NULL = (1 << 64) - 1
UNIVERSES = [1, 2, 63, 64, 65, 127, 128, 129, 512, 4095, 4096, 4097, 12800, 262145]


def layers_for(u):
    return (u.bit_length() - 1) // 6 + 1


def summary_consistent(t):
    L = t.layers()
    for layer in range(L - 1):
        up, low = t.words(layer), t.words(layer + 1)
        for w, val in enumerate(low):
            bit = (int(up[w // 64]) >> (w % 64)) & 1
            if bit != (1 if val else 0):
                return False
    return True


@pytest.mark.parametrize("u", UNIVERSES)
def test_layers_and_initial_state(u):
    full = oracle.Veb(u, full=True)
    assert full.layers() == layers_for(u)
    bottom = full.words(full.layers() - 1)
    assert sum(bin(int(x)).count("1") for x in bottom) == u
    assert summary_consistent(full)
    empty = oracle.Veb(u)
    assert all(int(x) == 0 for layer in range(empty.layers()) for x in empty.words(layer))


@pytest.mark.parametrize("u", UNIVERSES)
def test_insert_remove_successor_against_set(u):
    rng = random.Random(u)
    t = oracle.Veb(u)
    model = set()
    for step in range(min(4000, 40 * u + 200)):
        v = rng.randrange(u)
        if rng.random() < 0.55:
            assert t.insert(v) == (v not in model)
            model.add(v)
        else:
            assert t.remove(v) == (v in model)
            model.discard(v)
        if step % 7 == 0:
            q = rng.randrange(u)
            assert t.query(q) == (q in model)
            above = [y for y in model if y > q]
            assert t.successor(q) == (min(above) if above else NULL)
    assert summary_consistent(t)


@pytest.mark.parametrize("u", [64, 129, 4097, 12800])
def test_malloc_first_takes_lowest(u):
    rng = random.Random(3)
    t = oracle.Veb(u)
    model = set(rng.sample(range(u), min(u, 50)))
    for v in model:
        t.insert(v)
    while model:
        lo = min(model)
        assert t.malloc_first() == lo
        model.remove(lo)
        assert not t.query(lo)
    assert t.malloc_first() == NULL


def expected_find_random(model, u, seed, tid):
    key = struct.pack("<Q", tid)
    h1 = py_murmur(key, seed)
    h2 = py_murmur(key, h1)
    for a in range(15):
        s = (h1 + a * h2) % (1 << 64) % u
        if s in model:
            return s
        above = [y for y in model if y > s]
        if above:
            return min(above)
    above = [y for y in model if y > 0]
    return min(above) if above else NULL


@pytest.mark.parametrize("u,seed", [(64, 111), (512, 112), (4096, 5), (12800, 119)])
def test_find_random_valid_index(u, seed):
    rng = random.Random(u)
    for fill in (0, 1, 2, 5, u // 10, u // 2):
        t = oracle.Veb(u, seed=seed)
        model = set(rng.sample(range(u), fill))
        for v in model:
            t.insert(v)
        for tid in [0, 1, 2, 63, 64, 1000, 123456, (1 << 32) + 5]:
            assert t.find_random(tid) == expected_find_random(model, u, seed, tid)
    # the lowest item is missed when it is alone and no hashed start hits it
    t = oracle.Veb(u, seed=seed)
    t.insert(0)
    for tid in range(20):
        exp = expected_find_random({0}, u, seed, tid)
        assert exp in (0, NULL)
        assert t.find_random(tid) == exp


def bits_of_words(words):
    out = []
    for w, val in enumerate(int(x) for x in words):
        while val:
            low = val & -val
            out.append(w * 64 + low.bit_length() - 1)
            val ^= low
    return out


def trailing_ones(x):
    n = 0
    while x & 1:
        n += 1
        x >>= 1
    return n


def leading_ones(x):
    n = 0
    while n < 64 and (x >> (63 - n)) & 1:
        n += 1
    return n


def expected_gather(words, n):
    W = len(words)
    for w in range(W - 1, 0, -1):
        t = trailing_ones(words[w])
        if t == 0:
            continue
        if t >= n:
            return w * 64 + t - n
        need = n - t
        k = w - 1
        ok = True
        while need > 64:
            if k < 0 or words[k] != (1 << 64) - 1:
                ok = False
                break
            need -= 64
            k -= 1
        if ok and k >= 0 and leading_ones(words[k]) >= need:
            return w * 64 + t - n
    return NULL


@pytest.mark.parametrize("u", [64, 128, 256, 320, 4096])
def test_gather_multiple_rule(u):
    rng = random.Random(u * 7 + 1)
    W = (u - 1) // 64 + 1
    for trial in range(600):
        # runs of ones with random holes, biased to produce long runs
        bits = 0
        p_one = rng.choice([0.5, 0.9, 0.98, 1.0])
        for i in range(u):
            if rng.random() < p_one:
                bits |= 1 << i
        words = [(bits >> (64 * w)) & ((1 << 64) - 1) for w in range(W)]
        t = oracle.Veb(u)
        t.set_words(t.layers() - 1, np.array(words, dtype=np.uint64))
        n = rng.choice([1, 2, 3, 5, 17, 63, 64, 65, 100, 128, 129, 200, 300])
        exp = expected_gather(words, n)
        got = t.gather_multiple(n)
        after = [int(x) for x in t.words(t.layers() - 1)]
        assert got == exp, (trial, n, [hex(x) for x in words])
        if got == NULL:
            assert after == words  # a failed gather changes nothing
        else:
            claimed = ((1 << n) - 1) << got
            before_bits = sum(w << (64 * i) for i, w in enumerate(words))
            after_bits = sum(w << (64 * i) for i, w in enumerate(after))
            assert before_bits & claimed == claimed
            assert after_bits == before_bits & ~claimed


def test_gather_never_served_from_a_single_word():
    t = oracle.Veb(64, full=True)
    assert t.gather_multiple(1) == NULL
    assert t.gather_multiple(2) == NULL


def test_return_multiple_sets_summary_only_for_full_words():
    u = 4096
    t = oracle.Veb(u)
    t.return_multiple(70, 100)  # words 1 (bits 6..63) and 2 (bits 0..41)
    bottom = [int(x) for x in t.words(t.layers() - 1)]
    expect = ((1 << 100) - 1) << 70
    assert sum(w << (64 * i) for i, w in enumerate(bottom)) == expect
    assert int(t.words(t.layers() - 2)[0]) == 0  # neither word is full
    t.return_multiple(128, 64)  # word 2 becomes full
    assert (int(t.words(t.layers() - 2)[0]) >> 2) & 1 == 1
    # gather then return round trip restores the item layer
    t2 = oracle.Veb(u, full=True)
    first = t2.gather_multiple(300)
    assert first != NULL
    t2.return_multiple(first, 300)
    assert all(int(x) == (1 << 64) - 1 for x in t2.words(t2.layers() - 1))
