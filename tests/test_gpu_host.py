"""Pools in mapped host memory, the combined and fused global allocators,
and the poison allocator (reference global_allocator.cuh:98-357,
tests/src/poison_test.cu), on the GPU."""
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


def test_host_pool_replay_matches_oracle_and_concurrent_kernel():
    pool, seed = 256 * MiB, 7
    g = ga.Gallatin.generate_on_device_host(pool, seed)
    try:
        o = oracle.Oracle(pool, seed)
        ops = [(0, 16 << (i % 9), i, i % 128) for i in range(2000)] + [(1, i, 0, 0) for i in range(0, 2000, 3)]
        assert np.array_equal(g.replay(ops), o.replay(ops)), "host-pool replay differs from the oracle"
        assert np.array_equal(g.snapshot(), o.snapshot())
        r = g.run_plain(1 << 16, 64, 3, verify=ga.GAL_VERIFY_ATOMIC)
        say(f"[host-pool] plain {r}")
        assert r["misses"] == 0 and r["tag_errors"] == 0 and r["alloc_errors"] == 0
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()


def test_fused_global_spills_to_host_and_frees_both():
    ga.init_global_allocator_combined(512 * MiB, 1 << 30, 9, print_info=False)
    try:
        dev, host = ga.global_handles()
        assert dev is not None and host is not None
        on_dev, on_host, misses, bad = ga.global_fused(1 << 17, 4096)  # 512 MiB: more than the device pool holds
        say(f"[fused] device {on_dev} host {on_host} misses {misses} bad {bad}")
        assert on_dev > 0 and on_host > 0 and misses == 0 and bad == 0
        for g in (dev, host):
            assert g.errors() == (0, 0)
            invariants.check_quiescent(g.snapshot(), 16 * MiB, 16)
    finally:
        ga.free_global_allocator_combined()
    assert ga.global_handles() == (None, None)


@pytest.mark.parametrize("size", [4 << k for k in range(16)])  # 4 B .. 128 KiB, as the reference
def test_poison_guards(size):
    ga.init_global_allocator(1 << 30, 41, print_info=False)
    try:
        codes = ga.global_poison(size)
        say(f"[poison] size {size}: {codes}")
        assert codes[0] == 0 and codes[6] == 0, "an intact allocation failed its check"
        assert all(c > 0 for c in codes[1:6]), "an overwritten guard went unnoticed"
        dev, _ = ga.global_handles()
        assert dev.errors() == (0, 0)
        invariants.check_quiescent(dev.snapshot(), 16 * MiB, 16)
    finally:
        ga.free_global_allocator()
