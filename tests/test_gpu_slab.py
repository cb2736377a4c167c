"""The slab hash client on the GPU (gal_slab_*), and a small real autotune.

Every inserted key must be found with its value; the table must hold
between keys/7 slabs (all full) and the number of slab publications; the
allocator must report no internal error, and after the table is destroyed
every slab is back and the allocator is quiescent.
"""
import sys

import pytest

import gallatin_amd as ga
from gallatin_amd import autotune
from tests import invariants

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("triple", [(16 * MiB, 16, 4096), (4 * MiB, 128, 1024), (64 * MiB, 128, 4096)])
def test_slab_hash_build_query_destroy(triple):
    seg, mn, mx = triple
    pool = 4 << 30
    n, buckets = 1 << 21, 1 << 15
    g = ga.Gallatin.generate_on_device(pool, 31, segment_size=seg, min_size=mn, max_size=mx)
    try:
        t = ga.SlabTable(g, buckets, seed=77)
        r = t.insert(n, key_seed=5)
        say(f"[slab] {triple} insert {r}")
        assert r["failed"] == 0 and r["alloc_errors"] == 0
        assert n / 7 <= r["slabs"] <= n / 7 + 2 * buckets
        q = t.query(n, key_seed=5)
        say(f"[slab] query {q}")
        assert q["failed"] == 0 and q["wrong"] == 0
        miss = t.query(1 << 16, key_seed=1 << 40)  # a disjoint key stream: absent
        assert miss["failed"] == 1 << 16
        freed = t.destroy()
        assert freed == r["slabs"]
        assert g.errors() == (0, 0), g.error_sites()
        invariants.check_quiescent(g.snapshot(), seg, mn)
    finally:
        g.free_on_device()


def test_slab_hash_two_builds_reuse_memory():
    g = ga.Gallatin.generate_on_device(2 << 30, 32)
    try:
        for rep in range(2):
            with ga.SlabTable(g, 1 << 14, seed=rep) as t:
                r = t.insert(1 << 20, key_seed=rep)
                q = t.query(1 << 20, key_seed=rep)
                assert r["failed"] == 0 and q["failed"] == 0 and q["wrong"] == 0
        assert g.errors() == (0, 0)
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
    finally:
        g.free_on_device()


def test_autotune_small_real_trial():
    out = autotune.tune(4 << 30, slab_keys=1 << 18, slab_buckets=1 << 12, batch_requests=1 << 18, replays=1,
                        widths=(16, 64), runs=(1, 16))
    say(f"[autotune] best {out['triple']} width {out['width']} run {out['run']} rate {out['rate']:.3e}")
    for tr in out["trials"]:
        say(f"  {tr}")
    assert out["triple"] in autotune.candidates(128)
    n_cands = len(autotune.candidates(128, (4 << 20, 16 << 20, 64 << 20)))
    assert out["rate"] > 0 and len(out["trials"]) == n_cands + min(2, n_cands)  # probes, then the grid of 2
