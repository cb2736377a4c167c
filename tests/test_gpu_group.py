"""GPU parity of multi-lane call sites: one wavefront, exact exec masks.

gal_group_replay runs each op with exactly its lanes active, so the team
claims (one add per class team, lane-order prefixes, write-off of straddled
slices, replacement on slice 4095) and the per-block free teams are
deterministic; the results and the whole metadata image must equal the
oracle's bit for bit. The spread cases mix in GAL_OP_MALLOC_SPREAD ops, the
claims malloc makes under load: a group's first pass starts at slot
hint % usable of its class (pinned or overflow), fills an empty one unless
the class is out of memory, and fails at once when it is and every pinned
slot holds 0.
"""
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from gallatin_amd import snapshot as snap
from oracle import oracle
from tests import group_streams, invariants

pytestmark = pytest.mark.gpu

MiB = 1 << 20
NULL = (1 << 64) - 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    # name, pool, (seg, min, max), slice sizes, large sizes, ops, share of spread malloc ops
    ("slices", 1 << 30, (16 * MiB, 16, 4096), [16, 32, 48, 64, 256, 1000, 2048, 4096], (), 1500, 0.0),
    ("one_class_full_waves", 1 << 30, (16 * MiB, 16, 4096), [16], (), 800, 0.0),
    ("big_slices", 1 << 30, (16 * MiB, 16, 4096), [16, 4096, 8192, 12000, 16384, 32768], (), 1200, 0.0),
    ("large", 4 << 30, (16 * MiB, 16, 4096), [16, 256, 4096, 8192], (40000, MiB, 16 * MiB, 40 * MiB), 800, 0.0),
    ("exhaustion", 64 * MiB, (MiB, 16, 256), [16, 32, 256, 1024, 4096], (65536, 2 * MiB), 1500, 0.0),
    ("min128", 1 << 30, (4 * MiB, 128, 1024), [16, 100, 128, 512, 1024, 3000], (200000,), 1000, 0.0),
    ("spread_slices", 1 << 30, (16 * MiB, 16, 4096), [16, 32, 48, 64, 256, 1000, 2048, 4096], (), 1500, 0.7),
    ("spread_big_slices", 2 << 30, (16 * MiB, 16, 4096), [64, 512, 4096, 8192, 16384, 32768], (), 1200, 0.7),
    ("spread_large", 4 << 30, (16 * MiB, 16, 4096), [16, 256, 4096], (40000, MiB, 40 * MiB), 800, 0.5),
    ("spread_exhaustion", 64 * MiB, (MiB, 16, 256), [16, 32, 256, 1024, 4096], (65536, 2 * MiB), 1500, 0.7),
    ("spread_small_pool_oom", 256 * MiB, (16 * MiB, 16, 4096), [1024, 2048, 4096], (), 1500, 0.8),
]


@pytest.mark.parametrize("name,pool,triple,sizes,large,n,spread_p", CASES, ids=[c[0] for c in CASES])
def test_group_replay_bit_exact(name, pool, triple, sizes, large, n, spread_p):
    seg, mn, mx = triple
    seed = 31
    o = oracle.Oracle(pool, seed, segment=seg, min_size=mn, max_size=mx)
    ops, results, live = group_streams.build(o, seed, n, sizes, large, seg=seg, mn=mn, mx=mx,
                                             large_p=0.08 if large else 0.0, spread_p=spread_p)
    expect = np.array(results, dtype=np.uint64)
    lanes = sum(bin(op[2]).count("1") for op in ops)
    say(f"[group] {name}: {len(ops)} ops, {lanes} lane calls, {len(live)} live at the end")
    snap_mid = o.snapshot()
    # then everything still live is freed, in call sites of up to 64 lanes
    # (segment runs alone)
    tail = []
    runs = [e for e in live if e[2]]
    rest = [e for e in live if not e[2]]
    for k in range(0, len(rest), 64):
        chunk = rest[k:k + 64]
        args, offs, mask = [0] * 64, [0] * 64, 0
        for lane, e in enumerate(chunk):
            args[lane], offs[lane] = e[0], e[1]
            mask |= 1 << lane
        assert o.group_free(mask, offs) == 0
        tail.append((1, 0, mask, args, [0] * 64))
    for e in runs:
        args = [0] * 64
        args[5] = e[0]
        assert o.free(e[1]) == 0
        tail.append((1, 0, 1 << 5, args, [0] * 64))
    snap_end = o.snapshot()
    invariants.check_quiescent(snap_end, seg, mn)
    for what, stream, image in (("stream", ops, snap_mid), ("stream + frees", ops + tail, snap_end)):
        g = ga.Gallatin.generate_on_device(pool, seed, segment_size=seg, min_size=mn, max_size=mx)
        try:
            got = g.group_replay(stream)
            exp = expect if len(stream) == len(ops) else np.vstack(
                [expect, np.array([[0 if (op[2] >> k) & 1 else NULL for k in range(64)] for op in tail],
                                  dtype=np.uint64).reshape(-1, 64)])
            bad = np.nonzero((got != exp).any(axis=1))[0]
            assert len(bad) == 0, (f"{what}: op {bad[0]} lanes {stream[bad[0]][2]:#x}: GPU {got[bad[0]].tolist()} "
                                   f"oracle {exp[bad[0]].tolist()}")
            gs = g.snapshot()
            assert np.array_equal(gs, image), f"{what}: {snap.first_difference(gs, image)}"
            assert g.errors() == (0, 0)
            say(f"[group] {name}: {what} bit-exact")
        finally:
            g.free_on_device()
