"""Two oracle rules pinned directly against what fixes them.

* The re-register rule of return_block (reference gallatin.cuh:1646) is a
  floating-point compare on unsigned operands:
      1.0 * slot / nb >= .1  &&  1.0 * (slot - 1) / nb < .1
  with slot and slot - 1 unsigned 32-bit (slot - 1 wraps at slot 0). The
  oracle (and the GPU) use an integer form; it must agree with the
  floating-point compare, evaluated here in IEEE double as the reference
  does, for every power-of-two nb up to 2^12 (a class's blocks per segment)
  and every slot a queue can hand out (and past the end).
* A gather (veb.cuh:1160-1196) clears the summary bit of every word it
  empties, on every upper layer (maybe_remove_layer_bits, veb.cuh:1100-1126),
  and touches no summary bit outside the words that hold the claimed items
  and their summary ancestors. Checked on three- and four-layer trees: from
  exact summaries the summaries stay exact; stale bits elsewhere survive.
"""
import random

import numpy as np
import pytest

from oracle import oracle
from tests.test_oracle_veb import expected_gather

NULL = (1 << 64) - 1


def float_rule(slot, nb):
    s = np.uint32(slot)
    s1 = np.uint32((slot - 1) & 0xFFFFFFFF)
    return bool(1.0 * float(s) / float(nb) >= 0.1 and 1.0 * float(s1) / float(nb) < 0.1)


@pytest.mark.parametrize("nb", [1 << k for k in range(13)])
def test_reregister_integer_form_equals_the_float_compare(nb):
    hits = 0
    for slot in range(0, 4 * nb + 3):
        want = float_rule(slot, nb)
        assert bool(oracle.lib.orc_reregister(slot, nb)) == want, (slot, nb)
        hits += want
    # it fires for exactly one slot, the first past a tenth of nb; for
    # nb == 1 that is slot 1, which a one-entry queue never hands out (its
    # slot is always 0), so such a segment is never re-listed early
    assert hits == 1
    for slot in (0xFFFFFFFF, 0x80000000):
        assert bool(oracle.lib.orc_reregister(slot, nb)) == float_rule(slot, nb)


def summaries_exact(layers):
    for upper, lower in zip(layers[:-1], layers[1:]):
        for w, val in enumerate(lower):
            if ((int(upper[w // 64]) >> (w % 64)) & 1) != (1 if val else 0):
                return False
    return True


def tree_with(u, bottom, stale=()):
    """A tree over u items with the given item words and exact summaries,
    plus stale summary bits (layer, bit) set over empty words."""
    t = oracle.Veb(u)
    L = t.layers()
    layers = [None] * L
    layers[L - 1] = [int(x) for x in bottom]
    for layer in range(L - 2, -1, -1):
        lower = layers[layer + 1]
        words = [0] * len(t.words(layer))
        for w, val in enumerate(lower):
            if val:
                words[w // 64] |= 1 << (w % 64)
        layers[layer] = words
    for layer, bit in stale:
        layers[layer][bit // 64] |= 1 << (bit % 64)
    for layer in range(L):
        t.set_words(layer, np.array(layers[layer], dtype=np.uint64))
    return t, layers


def ancestors(first, n, L):
    """(layer, word) pairs that hold the claimed items or a summary bit above them."""
    out = set()
    lo, hi = first // 64, (first + n - 1) // 64
    for layer in range(L - 1, -1, -1):
        for w in range(lo, hi + 1):
            out.add((layer, w))
        lo, hi = lo // 64, hi // 64
    return out


@pytest.mark.parametrize("u", [12800, 262145])
def test_gather_clears_exactly_the_summaries_of_emptied_words(u):
    rng = random.Random(u)
    W = (u - 1) // 64 + 1
    for trial in range(60):
        # long runs of free items with holes, and whole empty stretches so
        # that some summary words hold zeros
        bottom = []
        for w in range(W):
            r = rng.random()
            if r < 0.3:
                bottom.append(0)
            elif r < 0.8:
                bottom.append((1 << 64) - 1)
            else:
                bottom.append(rng.getrandbits(64))
        hi_item = u - 64 * (W - 1)
        bottom[-1] &= (1 << hi_item) - 1 if hi_item < 64 else (1 << 64) - 1
        t, _ = tree_with(u, bottom)
        L = t.layers()
        # stale summary bits over empty item words (as return_multiple leaves)
        empties = [w for w, v in enumerate(bottom) if v == 0]
        stale = [(L - 2, w) for w in rng.sample(empties, min(3, len(empties)))]
        t, before = tree_with(u, bottom, stale)
        n = rng.choice([2, 63, 64, 65, 130, 500, 1000])
        exp = expected_gather(bottom, n)
        got = t.gather_multiple(n)
        assert got == exp
        after = [[int(x) for x in t.words(layer)] for layer in range(L)]
        if got == NULL:
            assert after == before
            continue
        touched = ancestors(got, n, L)
        # exact where the gather looked: a summary bit is set iff its word holds items
        for layer in range(L - 1):
            for w, val in enumerate(after[layer + 1]):
                if (layer + 1, w) in touched:
                    bit = (after[layer][w // 64] >> (w % 64)) & 1
                    assert bit == (1 if val else 0), (trial, layer, w)
        # untouched summary words keep every bit, stale ones included
        for layer in range(L - 1):
            for w in range(len(after[layer])):
                below = range(w * 64, min(w * 64 + 64, len(after[layer + 1])))
                if not any((layer + 1, b) in touched for b in below):
                    assert after[layer][w] == before[layer][w], (trial, layer, w)
        if not stale:
            assert summaries_exact(after)
