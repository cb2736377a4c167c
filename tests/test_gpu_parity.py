"""GPU parity: the HIP allocator against the CPU oracle, through the C-ABI.

Single-lane replay: the GPU must return bit-identical byte offsets AND leave
bit-identical metadata (every vEB word, chunk id, block counter, queue entry,
segment counter and pinned slot) after the same op stream.
Concurrent runs: exact invariants (no overlap, alignment, in-pool, no tag
clash, no internal error, quiescent metadata consistent after full drain).
"""
import random
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from gallatin_amd import snapshot as snap
from oracle import oracle
from tests import invariants

pytestmark = pytest.mark.gpu

NULL = (1 << 64) - 1
MiB = 1 << 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def valid_stream(o, ops):
    """Run ops on the oracle and keep only frees of mallocs that succeeded."""
    out, remap, res = [], {}, []
    for i, (kind, arg, tid, slot) in enumerate(ops):
        if kind == 0:
            remap[i] = len(out)
            out.append((0, arg, tid, slot))
            res.append(o.malloc(arg, tid, slot))
        else:
            j = remap[arg]
            if res[j] == NULL:
                continue
            assert o.free(res[j]) == 0
            out.append((1, j, tid, slot))
            res.append(0)
    return out, np.array(res, dtype=np.uint64)


def random_ops(seed, n, sizes, free_p=0.45, slots=256):
    rng = random.Random(seed)
    live, ops = [], []
    for _ in range(n):
        if live and rng.random() < free_p:
            ops.append((1, live.pop(rng.randrange(len(live))), 0, 0))
        else:
            ops.append((0, rng.choice(sizes), rng.randrange(1 << 24), rng.randrange(slots)))
            live.append(len(ops) - 1)
    return ops


CASES = [
    # name, pool, (seg, min, max), sizes, n_ops, free_p
    ("plain16", 1 << 30, (16 * MiB, 16, 4096), [16], 9000, 0.3),
    ("mixed_slices", 1 << 30, (16 * MiB, 16, 4096), [16, 32, 48, 64, 256, 1000, 2048, 4096], 8000, 0.45),
    ("big_slices", 1 << 30, (16 * MiB, 16, 4096), [16, 4097, 8192, 12000, 16384, 32768], 6000, 0.45),
    ("blocks_segments", 4 << 30, (16 * MiB, 16, 4096), [16, 40000, 64 * 1024, MiB, 16 * MiB, 17 * MiB, 40 * MiB],
     3000, 0.45),
    ("exhaustion", 64 * MiB, (MiB, 16, 256), [16, 32, 256, 1024, 8192, 65536, 2 * MiB], 8000, 0.35),
    ("min128", 1 << 30, (4 * MiB, 128, 1024), [16, 100, 128, 512, 1024, 3000, 200000], 6000, 0.45),
    # 4096 segments: the segment tree and every class tree have three layers
    # (the cases above have one or two), with segment runs that span item words
    ("three_layers", 64 << 30, (16 * MiB, 16, 4096),
     [16, 4096, 40000, 64 * 1024, MiB, 16 * MiB, 40 * MiB, 200 * MiB, 1100 * MiB], 3000, 0.45),
    # the slab-hash workload's autotuned instance, and the 16 KiB-slice one
    ("seg64_min128", 32 << 30, (64 * MiB, 128, 4096),
     [16, 100, 128, 1000, 4096, 8192, 65536, MiB, 100 * MiB], 4000, 0.45),
    ("seg64_max16k", 16 << 30, (64 * MiB, 16, 16384),
     [16, 48, 1000, 16384, 20000, 70000, 5 * MiB, 70 * MiB], 4000, 0.45),
]


@pytest.mark.parametrize("name,pool,cfg,sizes,n,free_p", CASES, ids=[c[0] for c in CASES])
def test_replay_bit_exact(name, pool, cfg, sizes, n, free_p):
    seg, mn, mx = cfg
    seed = 1000 + len(name)
    say(f"[replay] case={name} pool={pool} cfg={cfg} ops={n}")
    o = oracle.Oracle(pool, seed, segment=seg, min_size=mn, max_size=mx)
    g = ga.Gallatin.generate_on_device(pool, seed, segment_size=seg, min_size=mn, max_size=mx)
    try:
        # construction (the booted pinned blocks) is already bit-identical
        boot_g, boot_o = g.snapshot(), o.snapshot()
        assert snap.first_difference(boot_g, boot_o) is None
        assert np.array_equal(boot_g, boot_o)

        ops, expect = valid_stream(o, random_ops(seed, n, sizes, free_p))
        got = g.replay(ops)
        diff = np.nonzero(got != expect)[0]
        assert len(diff) == 0, f"first mismatch at op {diff[0]}: {ops[diff[0]]} gpu={got[diff[0]]} cpu={expect[diff[0]]}"
        sg, so = g.snapshot(), o.snapshot()
        assert snap.first_difference(sg, so) is None
        assert np.array_equal(sg, so)
        assert g.errors() == (0, 0)
        assert o.faults() == 0
    finally:
        g.free_on_device()


def test_replay_free_all_restores_quiescent_state():
    pool, seed = 1 << 30, 7
    o = oracle.Oracle(pool, seed)
    g = ga.Gallatin.generate_on_device(pool, seed)
    try:
        ops = [(0, 16 << (i % 9), i, i % 200) for i in range(20000)]
        ops += [(1, i, 0, 0) for i in range(20000)]
        ops, expect = valid_stream(o, ops)
        got = g.replay(ops)
        assert np.array_equal(got, expect)
        sg = g.snapshot()
        assert np.array_equal(sg, o.snapshot())
        invariants.check_quiescent(sg, 16 * MiB, 16)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("threads", [1 << 16, 1 << 20])
def test_plain_kernel_concurrent(threads):
    say(f"[plain] threads={threads}")
    g = ga.Gallatin.generate_on_device(1 << 30, 111)
    try:
        r = g.run_plain(threads, 16, 5, verify=ga.GAL_VERIFY_ATOMIC)
        assert r["misses"] == 0 and r["tag_errors"] == 0 and r["alloc_errors"] == 0
        assert r["mallocs"] == threads and r["frees"] == threads
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        # again, on the state the first run left
        r = g.run_plain(threads, 16, 6, verify=ga.GAL_VERIFY_ATOMIC)
        assert r["misses"] == 0 and r["tag_errors"] == 0 and r["alloc_errors"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("size", [24, 100, 4096, 8192, 32768, 1 << 20])
def test_plain_kernel_sizes(size):
    say(f"[plain-size] size={size}")
    g = ga.Gallatin.generate_on_device(4 << 30, 3)
    try:
        # every thread holds its allocation at once, and blocks are reused
        # only when all their slices are back: keep the peak within the pool
        granted = max(16, 1 << (size - 1).bit_length())
        threads = min(1 << 18, (1 << 30) // granted)
        r = g.run_plain(threads, size, 9, verify=ga.GAL_VERIFY_ATOMIC)
        say(f"  result {r} errors {g.errors()} sites {g.error_sites()}")
        assert r["tag_errors"] == 0 and r["alloc_errors"] == 0
        assert r["mallocs"] + r["misses"] == threads
        if size <= 32768:
            assert r["misses"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
    finally:
        g.free_on_device()


SPREAD_CASES = [
    # (segment, min, max), pool, size: the 1 GiB pool gives the 128 B, 512 B
    # and 2 KiB classes 64, 32 and 8 usable slots (more than their pinned
    # ones); on the 64 MiB pool the budget leaves every class at its pinned
    # slots
    ((16 * MiB, 16, 4096), 1 << 30, 128),
    ((16 * MiB, 16, 4096), 1 << 30, 512),
    ((16 * MiB, 16, 4096), 1 << 30, 2048),
    ((MiB, 16, 256), 64 * MiB, 64),
    # 64 MiB segments: the slab-hash instance's 256 B class, and a top class
    # whose block is a whole segment (16 KiB x 4096)
    ((64 * MiB, 128, 4096), 2 << 30, 256),
    ((64 * MiB, 16, 16384), 16 << 30, 16384),
]


@pytest.mark.parametrize("cfg,pool,size", SPREAD_CASES, ids=[f"{c[1] >> 20}MiB-{c[2]}B" for c in SPREAD_CASES])
def test_spread_claims_leave_nothing_stranded(cfg, pool, size):
    """Under load a class claims from up to its usable slots (pinned and
    overflow, within 1/16 of the pool in blocks): repeated kernels of 2^18
    concurrent malloc+free pairs never miss, clash or leak, and leave the
    metadata quiescent after each run."""
    seg, mn, mx = cfg
    say(f"[spread] pool={pool} cfg={cfg} size={size}")
    g = ga.Gallatin.generate_on_device(pool, 5, segment_size=seg, min_size=mn, max_size=mx)
    try:
        for rep in range(3):
            r = g.run_plain(1 << 18, size, 40 + rep, verify=ga.GAL_VERIFY_ATOMIC)
            assert r["mallocs"] == 1 << 18 and r["misses"] == 0, r
            assert r["tag_errors"] == 0 and r["alloc_errors"] == 0, r
            invariants.check_quiescent(g.snapshot(), seg, mn)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("verify", [ga.GAL_VERIFY_ATOMIC, ga.GAL_VERIFY_STORE])
def test_mixed_rounds_invariants(verify):
    """Atomic mode: tags exchanged in and out (sharing detected), frees with
    free(). Store mode: the free kernel reads the tag and frees with
    free_released (the tags were stored by the previous kernel)."""
    pool, threads, rounds = 2 << 30, 1 << 18, 6
    say(f"[mixed] pool={pool} threads={threads} rounds={rounds} verify={verify}")
    g = ga.Gallatin.generate_on_device(pool, 21)
    try:
        g.mixed_reset(threads, rounds, 77)
        for r in range(rounds):
            res = g.mixed_round(r, verify=verify)
            offs, sizes, slots = g.mixed_live()
            say(f"  round {r}: {res} live {len(offs)} sites {g.error_sites()}")
            invariants.check_live(offs, sizes, pool, 16 * MiB, 16, 4096, oracle.route)
            assert res["tag_errors"] == 0 and res["alloc_errors"] == 0, (res, g.error_sites())
        res = g.mixed_drain(verify=verify)
        assert res["tag_errors"] == 0 and res["alloc_errors"] == 0
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        lat = g.latency_ns()
        assert len(lat) > 0 and np.all(lat >= 0)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


def test_mixed_full_workload_repeatable():
    pool, threads, rounds = 8 << 30, 1 << 20, 4
    say(f"[mixed-full] pool={pool} threads={threads} rounds={rounds}")
    g = ga.Gallatin.generate_on_device(pool, 2)
    try:
        for step in range(2):
            res = g.run_mixed(threads, rounds, 100 + step, verify=ga.GAL_VERIFY_ATOMIC)
            assert res["tag_errors"] == 0 and res["alloc_errors"] == 0, (res, g.error_sites())
            assert res["frees"] == res["mallocs"]
            invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        # latency samples: one per wave and round of the last run, or of the
        # runs kept (gal_latency_keep; runs past the count record none)
        per_run = rounds * threads // 64
        assert len(g.latency_ns()) == per_run
        g.latency_keep(2)
        assert len(g.latency_ns()) == 0
        for step in range(3):
            res = g.run_mixed(threads, rounds, 200 + step)
            assert res["tag_errors"] == 0 and res["alloc_errors"] == 0
            lat = g.latency_ns()
            assert len(lat) == per_run * min(step + 1, 2) and np.all(lat > 0)
        g.latency_keep(0)
        g.run_mixed(threads, rounds, 300)
        assert len(g.latency_ns()) == per_run
    finally:
        g.free_on_device()


def test_global_allocator():
    say("[global]")
    ga.init_global_allocator(1 << 30, 111, print_info=False)
    try:
        misses, bad = ga.global_smoke(1 << 18, 64)
        assert misses == 0 and bad == 0
        ga.print_global_stats()
    finally:
        ga.free_global_allocator()


def test_mixed_reset_frees_an_unfinished_run():
    """A reset in the middle of a mixed run frees what that run left live:
    the metadata is quiescent right after it, and a new run starts clean."""
    pool, threads = 2 << 30, 1 << 18
    say(f"[mixed-reset] pool={pool} threads={threads}")
    g = ga.Gallatin.generate_on_device(pool, 23)
    try:
        g.mixed_reset(threads, 6, 91)
        for r in range(3):
            res = g.mixed_round(r)
            assert res["tag_errors"] == 0 and res["alloc_errors"] == 0, res
        assert len(g.mixed_live()[0]) > 0
        g.mixed_reset(threads, 2, 92)
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        for r in range(2):
            res = g.mixed_round(r)
            assert res["tag_errors"] == 0 and res["alloc_errors"] == 0, res
        g.mixed_drain()
        invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()
