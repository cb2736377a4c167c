"""gal_gather_limit (host code of the C-ABI library): the largest number of
contiguous segments the segment tree's gather rule could claim from given
item words. Pinned against the oracle's own gather_multiple (a literal
restatement of reference veb.cuh:323-452 with fix Q4), run on a fresh copy
of the same words for every n, and against hand-worked cases. No GPU.
"""
import numpy as np
import pytest

import gallatin_amd as ga
from oracle import oracle

FULL = (1 << 64) - 1


def oracle_limit(words):
    words = np.asarray(words, dtype=np.uint64)
    universe = len(words) * 64
    best = 0
    for n in range(1, universe + 1):
        t = oracle.Veb(universe)
        t.set_words(t.layers() - 1, words)
        if t.gather_multiple(n) != oracle.NULL:
            best = n
    return best


@pytest.mark.parametrize(
    "words, expect",
    [
        ([FULL], 0),  # one word: the rule never starts in word 0
        ([0, 0b111], 3),  # trailing ones of the top word
        ([0, 0b1110], 0),  # ones not touching bit 0 of their word are never a run's start
        ([FULL << 60 & FULL, 0b1], 5),  # 1 bit + the top 4 bits of word 0
        ([FULL, FULL, 0b11], 130),  # 2 bits + two full words below
        ([FULL, 0, FULL], 64),  # a hole stops the extension
        ([0, FULL, FULL], 128),  # may reach word 1 but not below word 0's top bits
    ],
)
def test_known_cases(words, expect):
    words = np.array(words, dtype=np.uint64)
    assert ga.gather_limit(words) == expect
    assert oracle_limit(words) == expect


@pytest.mark.parametrize("seed", range(12))
def test_matches_oracle_on_random_trees(seed):
    rng = np.random.default_rng(seed)
    nwords = int(rng.integers(2, 7))
    words = []
    for _ in range(nwords):
        kind = rng.integers(0, 4)
        if kind == 0:
            words.append(FULL)
        elif kind == 1:
            words.append(0)
        elif kind == 2:  # low run of ones
            words.append((1 << int(rng.integers(1, 64))) - 1)
        else:  # random bits, often with a high run
            words.append(int(rng.integers(0, 1 << 63)) | (FULL << int(rng.integers(1, 64)) & FULL))
    words = np.array(words, dtype=np.uint64)
    assert ga.gather_limit(words) == oracle_limit(words)
