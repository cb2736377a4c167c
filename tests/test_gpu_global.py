"""The global API on the GPU: init_global_allocator / global_malloc /
global_free / print_global_stats (reference global_allocator.cuh:41-97).

* A single lane through global_malloc/global_free returns the oracle's
  offsets and leaves the oracle's metadata: global_malloc passes its CU's
  slot hint (the CU id with its XCD in the low three bits) and thread id 0
  to the hashed search, so the oracle replays the stream with that slot and
  tid 0.
* The reference's global test kernel shape at its full size (1 GiB pool,
  2^20 threads x global_malloc(16 B), tag, check, global_free): no miss, no
  clash, quiescent metadata after.
* Handles that batches, serving loops or slab tables still use are not
  destroyed: the call is refused and the handle keeps working.
"""
import random
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from gallatin_amd import _capi
from gallatin_amd import snapshot as snap
from oracle import oracle
from tests import invariants

pytestmark = pytest.mark.gpu

MiB = 1 << 20
NULL = (1 << 64) - 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def global_stream(seed, n, sizes, free_p):
    rng = random.Random(seed)
    o = oracle.Oracle(1 << 30, seed)  # only tells which mallocs succeed
    ops, res, live = [], [], []
    for _ in range(n):
        if live and rng.random() < free_p:
            j = live.pop(rng.randrange(len(live)))
            ops.append((1, j, 0, 0))
            res.append(0)
        else:
            size = rng.choice(sizes)
            ops.append((0, size, 0, 0))
            res.append(o.malloc(size, 0, 0))
            if res[-1] != NULL:
                live.append(len(ops) - 1)
    return ops


@pytest.mark.parametrize("sizes", [[16], [16, 48, 256, 1000, 4096, 8192, 32768, 40000, MiB, 20 * MiB]],
                         ids=["16B", "mixed"])
def test_global_replay_bit_exact(sizes):
    seed = 57
    ops = global_stream(seed, 6000, sizes, 0.4)
    # the CU the lane will run on is known only afterwards; the stream is
    # valid (every free names a malloc that succeeded) for every slot hint
    # modulo 128, the largest slot count (replay raises otherwise)
    for slot in range(128):
        oracle.Oracle(1 << 30, seed).replay([(k, a, 0, slot) for (k, a, _, _) in ops])
    ga.init_global_allocator(1 << 30, seed, print_info=False)
    try:
        dev, _ = ga.global_handles()
        got, cu = ga.global_replay(ops)
        say(f"[global-replay] {len(ops)} ops with CU slot hint {cu}")
        # the same stream through the oracle, with the slot the lane ran on
        o = oracle.Oracle(1 << 30, seed)
        expect = o.replay([(k, a, 0, cu) for (k, a, _, _) in ops])
        diff = np.nonzero(got != expect)[0]
        assert len(diff) == 0, f"op {diff[0]} {ops[diff[0]]}: GPU {got[diff[0]]} oracle {expect[diff[0]]}"
        gs, os_ = dev.snapshot(), o.snapshot()
        assert np.array_equal(gs, os_), snap.first_difference(gs, os_)
        assert dev.errors() == (0, 0)
        ga.print_global_stats()
    finally:
        ga.free_global_allocator()


@pytest.mark.parametrize("verify", [ga.GAL_VERIFY_STORE, ga.GAL_VERIFY_ATOMIC])
def test_global_plain_full_size(verify):
    """1 GiB pool; one kernel of 2^20 threads, each global_malloc(16 B), a
    tag drawn from (seed, thread) written and checked, global_free."""
    ga.init_global_allocator(1 << 30, 3, print_info=False)
    try:
        dev, _ = ga.global_handles()
        for rep in range(3):
            r = ga.global_plain(1 << 20, 16, 100 + rep, verify=verify)
            say(f"[global-plain] rep {rep}: {r}")
            assert r["mallocs"] == 1 << 20 and r["frees"] == 1 << 20, r
            assert r["misses"] == 0 and r["tag_errors"] == 0 and r["alloc_errors"] == 0, r
            invariants.check_quiescent(dev.snapshot(), 16 * MiB, 16)
        assert dev.errors() == (0, 0)
    finally:
        ga.free_global_allocator()


def test_destroy_refused_while_a_batch_uses_the_handle():
    g = ga.Gallatin.generate_on_device(1 << 30, 5)
    b = ga.Batch(g, 1 << 12)
    try:
        assert ga._lib.gal_destroy(g._h) == _capi.GAL_ERR_STATE
        # the handle and the batch still work
        r = g.run_plain(1 << 14, 64, 1, verify=ga.GAL_VERIFY_ATOMIC)
        assert r["misses"] == 0 and r["tag_errors"] == 0
        b.generate(1 << 12, seed=1, size_classes=9)
        r = b.run(ga.FORM_BATCHED, width=64, run=4)
        assert r["misses"] == 0 and r["tag_errors"] == 0
    finally:
        # the wrapper releases the batch before it destroys the handle
        g.free_on_device()
    assert b._b is None


def test_global_free_refused_while_a_slab_table_uses_it():
    ga.init_global_allocator(1 << 30, 6, print_info=False)
    dev, _ = ga.global_handles()
    t = ga.SlabTable(dev, 1 << 10)
    try:
        assert ga._lib.gal_free_global_allocator() == _capi.GAL_ERR_STATE
        ins = t.insert(1 << 14, key_seed=3)
        assert ins["failed"] == 0
        r = ga.global_plain(1 << 14, 16, 1)
        assert r["misses"] == 0 and r["tag_errors"] == 0
    finally:
        t.destroy()
        ga.free_global_allocator()
    assert ga.global_handles() == (None, None)
