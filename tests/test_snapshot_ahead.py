"""The ahead words of a metadata snapshot (the blocks fetched ahead for the
slots, which only claims under load fetch) and the quiescence check's rules
for them, on images of the oracle (CPU only).

The oracle models no fetch ahead, so its ahead words are all 0, as the
GPU's are after single-lane or group replays (the replay tests compare the
whole image). A quiescent image may hold a block in an ahead word: a fresh
block (no slice claimed) of the slot's class, held nowhere else. Each rule
is checked by moving blocks between a slot entry and an ahead word of an
oracle image.
"""
import numpy as np
import pytest

from gallatin_amd import snapshot as snap
from oracle import oracle
from tests import invariants

MiB = 1 << 20
SEG, MIN = 16 * MiB, 16


def slot_base(c):
    """first slot of class c: each class has max(pinned, 64) slots"""
    return sum(max(invariants.pinned_slots_of(k), 64) for k in range(c))


@pytest.fixture
def image():
    o = oracle.Oracle(1 << 30, 9)
    # class-0 claims from its pinned slot 0, then frees all: a quiescent image
    # whose pinned class-0 block in slot 0 has slices counted (and returned)
    offs = [o.malloc(16, tid=i, slot=0) for i in range(100)]
    for off in offs:
        assert o.free(off) == 0
    img = o.snapshot()
    invariants.check_quiescent(img, SEG, MIN)
    return np.array(img, dtype=np.uint8)


def test_oracle_images_list_zero_ahead_words(image):
    s = snap.parse(image)
    assert len(s["ahead"]) == len(s["slots"])
    assert not np.any(s["ahead"])


def test_fresh_block_in_ahead_word_is_quiescent(image):
    s = snap.parse(image)
    count = s["malloc_counter"].astype(np.int64) & ((1 << 20) - 1)
    # a pinned class-1 block (boot pins one per pinned slot, none claimed)
    b = slot_base(1)
    assert s["slots"][b] != 0 and count[s["slots"][b] - 1] == 0
    s["ahead"][b + 3] = s["slots"][b]  # the views write through to the image
    s["slots"][b] = 0
    invariants.check_quiescent(image, SEG, MIN)


def test_block_with_claimed_slices_in_ahead_word_is_rejected(image):
    s = snap.parse(image)
    count = s["malloc_counter"].astype(np.int64) & ((1 << 20) - 1)
    assert count[s["slots"][0] - 1] == 100
    s["ahead"][1] = s["slots"][0]
    s["slots"][0] = 0
    with pytest.raises(AssertionError, match="slices claimed"):
        invariants.check_quiescent(image, SEG, MIN)


def test_block_of_another_class_in_ahead_word_is_rejected(image):
    s = snap.parse(image)
    b1 = slot_base(1)
    s["ahead"][0] = s["slots"][b1]  # a class-1 block in a class-0 ahead word
    s["slots"][b1] = 0
    with pytest.raises(AssertionError, match="another class"):
        invariants.check_quiescent(image, SEG, MIN)


def test_block_held_twice_is_rejected(image):
    s = snap.parse(image)
    b = slot_base(1)
    s["ahead"][b + 1] = s["slots"][b]  # held by the slot and an ahead word
    with pytest.raises(AssertionError, match="pinned twice"):
        invariants.check_quiescent(image, SEG, MIN)


def test_unfinished_fetch_is_rejected(image):
    s = snap.parse(image)
    s["ahead"][5] = 0xFFFFFFFF
    with pytest.raises(AssertionError, match="unfinished"):
        invariants.check_quiescent(image, SEG, MIN)
