"""Batched service of uneven requests, through the C-ABI (gal_batch_*).

The binned order is checked exactly against a stable sort computed here; the
results of a batched malloc must satisfy the live invariants (no overlap,
alignment, inside the pool), the batched free must leave quiescent metadata,
and every form (per-thread, batched, hipGraph-replayed) must serve every
request with no tag clash and no internal error.
"""
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from oracle import oracle
from tests import invariants

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def expected_order(sizes, min_size):
    """Request indices sorted by bin (the claim class, capped at 31), stably."""
    sizes = np.asarray(sizes, dtype=np.uint64)
    ceil_log2 = np.array([(int(s) - 1).bit_length() for s in sizes.tolist()], dtype=np.int64)
    tree = np.where(sizes < min_size, 0, ceil_log2 - (min_size.bit_length() - 1))
    bins = np.minimum(tree, 31)
    return np.argsort(bins, kind="stable").astype(np.uint32)


def test_batch_binning_and_batched_malloc_free():
    # 4 GiB: the classes' segments (taken from the bottom of the segment tree)
    # leave the top words whole for the three segment runs
    pool, seed = 4 << 30, 5
    rng = np.random.default_rng(11)
    n = 70001  # not a multiple of the binning chunk
    sizes = (16 << rng.integers(0, 9, n)).astype(np.uint64)
    sizes[rng.integers(0, n, 2000)] = 8192  # big slices (2 x 4 KiB)
    sizes[rng.integers(0, n, 300)] = 60000  # big slices (16 x 4 KiB)
    sizes[rng.integers(0, n, 12)] = 1 << 20  # blocks
    sizes[rng.integers(0, n, 3)] = 40 << 20  # segment runs
    sizes[rng.integers(0, n, 500)] = rng.integers(1, 5000, 500).astype(np.uint64)  # odd sizes
    say(f"[batch-bin] n={n}")
    g = ga.Gallatin.generate_on_device(pool, seed)
    try:
        with ga.Batch(g, n) as b:
            b.set_sizes(sizes, seed=3)
            r = b.malloc(width=64)
            say(f"  malloc {r} errors {g.errors()}")
            assert np.array_equal(b.order(), expected_order(sizes, 16))
            offs = b.offsets()
            ok = offs != ga.GAL_NULL_OFFSET
            assert r["mallocs"] == int(ok.sum()) and r["misses"] == n - int(ok.sum())
            assert r["misses"] == 0
            invariants.check_live(offs[ok], sizes[ok], pool, 16 * MiB, 16, 4096, oracle.route)
            f = b.free()
            say(f"  free {f}")
            assert f["frees"] == r["mallocs"] and f["tag_errors"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("form", [ga.FORM_PER_THREAD, ga.FORM_BATCHED, ga.FORM_GRAPH], ids=["per_thread", "batched", "graph"])
def test_batch_forms_serve_every_request(form):
    pool, n = 4 << 30, 1 << 20
    say(f"[batch-form] form={form} n={n}")
    g = ga.Gallatin.generate_on_device(pool, 8)
    try:
        with ga.Batch(g, n) as b:
            b.generate(n, seed=21, size_classes=9)  # 16 B .. 4 KiB
            r = b.run(form, width=64, replays=3, verify=ga.GAL_VERIFY_ATOMIC)
            say(f"  {r}")
            assert r["mallocs"] == 3 * n and r["frees"] == 3 * n and r["misses"] == 0
            assert r["tag_errors"] == 0 and r["alloc_errors"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("run", [2, 5, 8, 16])
def test_batch_bulk_runs(run):
    """Lanes claim `run` requests at once (bulk claim, free_run); runs that mix
    classes at bin edges and tails shorter than `run` fall back per request."""
    pool, n = 4 << 30, (1 << 20) + 77
    say(f"[batch-run] run={run}")
    g = ga.Gallatin.generate_on_device(pool, 13)
    try:
        with ga.Batch(g, n) as b:
            b.generate(n, seed=run, size_classes=9)
            r = b.run(ga.FORM_GRAPH, width=64, run=run, replays=2, verify=ga.GAL_VERIFY_ATOMIC)
            say(f"  {r} sites {g.error_sites()}")
            assert r["mallocs"] == 2 * n and r["frees"] == 2 * n and r["misses"] == 0
            assert r["tag_errors"] == 0 and r["alloc_errors"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


def test_batch_bulk_malloc_then_single_frees():
    """Results of bulk claims are ordinary allocations: the live invariants hold
    and freeing them one by one restores quiescent metadata."""
    pool, seed = 2 << 30, 6
    rng = np.random.default_rng(3)
    n = 50003
    sizes = (16 << rng.integers(0, 9, n)).astype(np.uint64)
    sizes[rng.integers(0, n, 3000)] = 12000  # big slices (4 x 4 KiB)
    g = ga.Gallatin.generate_on_device(pool, seed)
    try:
        with ga.Batch(g, n) as b:
            b.set_sizes(sizes, seed=9)
            r = b.malloc(width=64, run=8)
            assert r["misses"] == 0 and r["mallocs"] == n
            offs = b.offsets()
            invariants.check_live(offs, sizes, pool, 16 * MiB, 16, 4096, oracle.route)
            f = b.free()
            assert f["frees"] == n and f["tag_errors"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("width", [1, 8, 16, 32, 64])
def test_batch_group_widths(width):
    pool, n = 2 << 30, 1 << 18
    say(f"[batch-width] width={width}")
    g = ga.Gallatin.generate_on_device(pool, 9)
    try:
        with ga.Batch(g, n) as b:
            b.generate(n, seed=width, size_classes=9)
            r = b.run(ga.FORM_BATCHED, width=width, replays=2, verify=ga.GAL_VERIFY_ATOMIC)
            assert r["mallocs"] == 2 * n and r["frees"] == 2 * n
            assert r["tag_errors"] == 0 and r["alloc_errors"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
    finally:
        g.free_on_device()


def test_batch_graph_replays_match_eager():
    """Graph replays separated by eager work and synchronisations each serve
    the whole batch, like the eager launches do."""
    pool, n = 2 << 30, 1 << 18
    g = ga.Gallatin.generate_on_device(pool, 10)
    try:
        with ga.Batch(g, n) as b:
            b.generate(n, seed=4, size_classes=9)
            eager = b.run(ga.FORM_BATCHED, replays=1, verify=ga.GAL_VERIFY_ATOMIC)
            for k in range(3):
                graph = b.run(ga.FORM_GRAPH, replays=2, verify=ga.GAL_VERIFY_ATOMIC)
                say(f"  replay set {k}: {graph}")
                for key in ("mallocs", "frees", "misses", "tag_errors", "alloc_errors"):
                    assert graph[key] == 2 * eager[key], (key, graph, eager)
                again = b.run(ga.FORM_BATCHED, replays=1, verify=ga.GAL_VERIFY_ATOMIC)
                assert again["mallocs"] == eager["mallocs"] and again["tag_errors"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("seg,mx", [(4 * MiB, 1024), (16 * MiB, 4096), (64 * MiB, 4096)])
def test_batch_min128_instantiations(seg, mx):
    """The slab-hash tuning space: min 128 B, sizes 16 B .. 64 KiB (big slices above max)."""
    pool, n = 8 << 30, 1 << 19
    say(f"[batch-min128] seg={seg} max={mx}")
    g = ga.Gallatin.generate_on_device(pool, 12, segment_size=seg, min_size=128, max_size=mx)
    try:
        with ga.Batch(g, n) as b:
            b.generate(n, seed=33, size_classes=13)  # 16 B .. 64 KiB
            r = b.run(ga.FORM_GRAPH, width=64, run=4, replays=2, verify=ga.GAL_VERIFY_ATOMIC)
            say(f"  {r} error sites {g.error_sites()}")
            assert r["mallocs"] + r["misses"] == 2 * n and r["frees"] == r["mallocs"]
            assert r["tag_errors"] == 0 and r["alloc_errors"] == 0
            assert r["misses"] == 0
        invariants.check_quiescent(g.snapshot(), seg, 128)
    finally:
        g.free_on_device()
