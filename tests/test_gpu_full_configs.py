"""The workloads at their full specified sizes, with their invariants.

* Mixed slice sizes: 8 GiB pool, 2^22 threads x 10 rounds, sizes 16 B-4 KiB
  log-uniform, seeded 50 % frees after each round; after every round the
  live allocations do not overlap, are aligned and lie in the pool, no tag
  was clobbered, no internal error; after the drain the metadata is
  quiescent.
* Batched: the slab hash (2^26 keys into 2^20 buckets on a 32 GiB pool of
  the tuned instance, every key queried), then 2^24 mixed 16 B-64 KiB
  requests x 100 hipGraph replays on the live table: no tag clobbered, no
  internal error, quiescent after the table is destroyed.
* Serving: 10^6 requests, Poisson 50k/s, lognormal lengths of 16 KiB pages
  on 200 GiB: every request finishes, no page retry, and the checksums of a
  seeded sample of 2 * 10^4 requests equal the host computation.
* The reference's mainline test shape (tests/src/gallatin_test.cu:58-219):
  90 % of an 8 GiB pool at one size (16 B, 4 KiB, 64 KiB) live at once,
  tags checked, all freed, repeated; no block hands out more than 4096
  allocations (tests/src/global_test_noprint.cu:101,143-166).
"""
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from oracle import oracle
from tests import invariants, serving_ref

pytestmark = pytest.mark.gpu

MiB = 1 << 20
GiB = 1 << 30


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_mixed_full_size_invariants_every_round():
    pool, threads, rounds = 8 * GiB, 1 << 22, 10
    g = ga.Gallatin.generate_on_device(pool, 2024)
    try:
        g.mixed_reset(threads, rounds, 2024)
        for r in range(rounds):
            res = g.mixed_round(r)
            offs, sizes, _ = g.mixed_live()
            say(f"[mixed-full] round {r}: mallocs {res['mallocs']} misses {res['misses']} frees {res['frees']} "
                f"live {len(offs)}")
            assert res["tag_errors"] == 0 and res["alloc_errors"] == 0, (res, g.error_sites())
            invariants.check_live(offs, sizes, pool, 16 * MiB, 16, 4096, oracle.route)
            # between kernels the fast out-of-memory test's count is the true one
            assert g.fill()["empty_pinned_slots"] == invariants.empty_pinned_counts(g.snapshot())
        res = g.mixed_drain()
        assert res["tag_errors"] == 0 and res["alloc_errors"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


def test_slab_and_graph_batches_full_size():
    pool, seg, mn, mx = 32 * GiB, 64 * MiB, 128, 4096
    keys, buckets = 1 << 26, 1 << 20
    g = ga.Gallatin.generate_on_device(pool, 9, segment_size=seg, min_size=mn, max_size=mx)
    try:
        t = ga.SlabTable(g, buckets, seed=9)
        ins = t.insert(keys, key_seed=9)
        say(f"[slab-full] insert {ins}")
        assert ins["failed"] == 0 and ins["alloc_errors"] == 0
        assert keys // 7 <= ins["slabs"] <= keys // 7 + 2 * buckets
        q = t.query(keys, key_seed=9)
        assert q["failed"] == 0 and q["wrong"] == 0, q
        with ga.Batch(g, 1 << 24) as b:
            b.generate(1 << 24, seed=9, size_classes=13)
            r = b.run(ga.FORM_GRAPH, width=64, run=32, replays=100)
            say(f"[slab-full] 100 graph replays: {r}")
            assert r["tag_errors"] == 0 and r["alloc_errors"] == 0, (r, g.error_sites())
            assert r["mallocs"] + r["misses"] == 100 * (1 << 24)
            assert r["frees"] == r["mallocs"]
        q = t.query(keys, key_seed=9)  # the table survived the batches
        assert q["failed"] == 0 and q["wrong"] == 0, q
        freed = t.destroy()
        assert freed == ins["slabs"]
        invariants.check_quiescent(g.snapshot(), seg, mn)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


def test_serving_full_size_checksums():
    reqs = ga.generate_requests(1_000_000, 50000.0, median_pages=1024, sigma=1.0, max_pages=32768,
                                page_bytes=16384, seed=2024)
    g = ga.Gallatin.generate_on_device(200 * GiB, 2024)
    try:
        with ga.Serving(g, len(reqs)) as s:
            s.load(reqs)
            r = s.run(slots=16384, run=64, deadline_s=float(reqs["arrival_ns"][-1]) * 1e-9 + 60.0)
            sums, admit, done = s.results()
        say(f"[serving-full] {r}")
        assert r["requests_done"] == len(reqs) and r["requests_aborted"] == 0
        assert r["page_retries"] == 0 and r["alloc_errors"] == 0
        pick = np.random.default_rng(7).choice(len(reqs), size=20000, replace=False)
        expect = serving_ref.checksums_c(reqs["id"][pick], reqs["pages"][pick], reqs["page_bytes"][pick])
        assert np.array_equal(sums[pick], expect)
        assert np.all(done[pick] >= admit[pick])
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("size,rounds", [(16, 2), (4096, 3), (65536, 3)], ids=["16B", "4KiB", "64KiB"])
def test_fill_ninety_percent_at_one_size(size, rounds):
    pool = 8 * GiB
    n = int(0.9 * pool) // size
    g = ga.Gallatin.generate_on_device(pool, 5)
    try:
        with ga.Batch(g, n) as b:
            if size == 16:
                b.generate(n, seed=5, size_classes=1)  # 16 << 0
            else:
                b.set_sizes(np.full(n, size, dtype=np.uint64), seed=5)
            granted, align = invariants.granted_bytes(16 * MiB, 16, 4096, size, oracle.route)
            block_bytes = size * 4096 if size <= 4096 else granted
            for rnd in range(rounds):
                m = b.malloc(width=64, run=1)
                say(f"[fill] {size} B round {rnd}: {m}")
                assert m["mallocs"] == n and m["misses"] == 0, m
                offs = b.offsets()
                assert np.all(offs != ga.GAL_NULL_OFFSET)
                assert np.all(offs % np.uint64(align) == 0) and np.all(offs + np.uint64(granted) <= np.uint64(pool))
                per_block = np.bincount((offs // np.uint64(block_bytes)).astype(np.int64))
                assert per_block.max() <= 4096
                f = b.free()
                assert f["frees"] == n and f["tag_errors"] == 0 and f["alloc_errors"] == 0, f
                invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()
