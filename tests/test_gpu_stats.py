"""The fill report (gal_fill: print_info's figures as numbers) on the GPU:
after a single-lane replay it must equal the figures counted from the
oracle's metadata snapshot, and after a concurrent kernel those counted from
the GPU's own snapshot."""
import numpy as np
import pytest

import gallatin_amd as ga
from gallatin_amd import snapshot as snap
from oracle import oracle
from tests import invariants

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def counted(image, seg, mn):
    s = snap.parse(image)
    T = s["num_trees"]
    chunk = s["chunk_ids"].astype(np.int64)
    active = s["active"].astype(np.int64)
    out = {
        "free_segments": len(invariants.bits_of(s["segment_tree"][-1])),
        "max_segments": s["num_segments"],
        "listed_segments": [len(invariants.bits_of(s["sub_trees"][c][-1])) for c in range(T)],
        "live_segments": [], "blocks_available": [], "blocks_in_use": [],
    }
    for c in range(T):
        nb = seg // ((mn << c) * 4096)
        live = (chunk == c) & (active != -1)
        out["live_segments"].append(int(live.sum()))
        out["blocks_available"].append(int(active[live].sum()))
        out["blocks_in_use"].append(int((nb - active[live]).sum()))
    return out


def test_fill_matches_oracle_after_replay():
    pool, seed = 512 * MiB, 4
    g = ga.Gallatin.generate_on_device(pool, seed)
    try:
        o = oracle.Oracle(pool, seed)
        ops = [(0, 16 << (i % 11), i, i % 64) for i in range(4000)] + [(1, i, 0, 0) for i in range(0, 4000, 2)]
        g.replay(ops)
        o.replay(ops)
        f = g.fill()
        want = counted(o.snapshot(), 16 * MiB, 16)
        for k, v in want.items():
            assert f[k] == v, (k, f[k], v)
        assert f["overhead_bytes"] == g.info.metadata_bytes > 0
    finally:
        g.free_on_device()


def test_fill_matches_snapshot_after_concurrent_kernel():
    g = ga.Gallatin.generate_on_device(2 << 30, 5, segment_size=64 * MiB, min_size=128, max_size=4096)
    try:
        g.run_plain(1 << 20, 1000, 3)
        f = g.fill()
        want = counted(g.snapshot(), 64 * MiB, 128)
        for k, v in want.items():
            assert f[k] == v, (k, f[k], v)
    finally:
        g.free_on_device()
