"""Host drivers of multi-round workloads built on the C-ABI.

block_churn: the block and multi-segment path. Each round is one concurrent
kernel (gal_churn) of interleaved mallocs and frees: an op is a malloc with
probability malloc_fraction, of floor(2^u) bytes with u uniform in
[lo_log2, hi_log2), else a free of an allocation live at the start of the
round (chosen without replacement). Sizes up to max_size * 4096 take whole
blocks, larger ones runs of segments (reference gallatin.cuh:790-861). After
every round the driver records the largest request the segment tree could
still satisfy (gal_largest_run): the fragmentation of the path.
"""
import time

import numpy as np

from ._capi import GAL_NULL_OFFSET, GAL_OP_FREE, GAL_OP_MALLOC


class BlockChurn:
    def __init__(self, allocator, requests=1 << 16, rounds=20, malloc_fraction=0.6, lo_log2=16, hi_log2=28, seed=0):
        self.g = allocator
        self.requests, self.rounds = requests, rounds
        self.malloc_fraction = malloc_fraction
        self.lo_log2, self.hi_log2 = lo_log2, hi_log2
        self.seed = seed
        self.rng = np.random.default_rng(seed)
        self.live = {}  # byte offset -> requested size
        self.round = 0
        self.history = []

    def round_ops(self):
        """(kinds, args, sizes) of the next round."""
        k = self.requests // self.rounds + (1 if self.round < self.requests % self.rounds else 0)
        is_malloc = self.rng.random(k) < self.malloc_fraction
        u = self.rng.uniform(self.lo_log2, self.hi_log2, k)
        sizes = np.floor(np.exp2(u)).astype(np.uint64)
        n_free = min(int((~is_malloc).sum()), len(self.live))
        offsets = np.array(sorted(self.live), dtype=np.uint64)
        victims = self.rng.choice(offsets, n_free, replace=False) if n_free else np.zeros(0, dtype=np.uint64)
        # frees beyond what is live are dropped; the rest keep their interleaved position
        free_pos = np.nonzero(~is_malloc)[0][:n_free]
        keep = is_malloc.copy()
        keep[free_pos] = True
        kinds = np.where(is_malloc, GAL_OP_MALLOC, GAL_OP_FREE).astype(np.uint32)[keep]
        args = np.where(is_malloc, sizes, np.uint64(0))
        args[free_pos] = victims
        return kinds, args[keep], sizes[keep]

    def step(self):
        kinds, args, sizes = self.round_ops()
        t0 = time.perf_counter()
        results, r = self.g.churn(kinds, args, seed=self.seed)
        wall = time.perf_counter() - t0
        # a malloc of the round may reuse memory a free of the same round gave
        # back: apply every free before recording the round's mallocs
        for kind, arg in zip(kinds.tolist(), args.tolist()):
            if kind == GAL_OP_FREE:
                del self.live[arg]
        for kind, size, res in zip(kinds.tolist(), sizes.tolist(), results.tolist()):
            if kind == GAL_OP_MALLOC and res != GAL_NULL_OFFSET:
                assert res not in self.live, f"offset {res} handed out twice"
                self.live[res] = size
        largest = self.g.largest_segment_run()
        rec = {
            "round": self.round,
            "mallocs": r["mallocs"],
            "frees": r["frees"],
            "misses": r["misses"],
            "tag_errors": r["tag_errors"],
            "alloc_errors": r["alloc_errors"],
            "kernel_ms": r["malloc_kernel_ms"],
            "wall_ms": wall * 1e3,
            "live": len(self.live),
            "live_bytes": int(sum(self.live.values())),
            "largest_run_segments": largest,
            "largest_request_bytes": largest * int(self.g.info.segment_size),
            "free_bytes": self.g.free_bytes(),
        }
        self.history.append(rec)
        self.round += 1
        return rec

    def run(self):
        while self.round < self.rounds:
            self.step()
        return self.history

    def live_arrays(self):
        offs = np.array(sorted(self.live), dtype=np.uint64)
        return offs, np.array([self.live[int(o)] for o in offs], dtype=np.uint64)

    def drain(self):
        """Free everything still live, in one concurrent kernel."""
        offs, _ = self.live_arrays()
        if len(offs) == 0:
            return None
        _, r = self.g.churn(np.full(len(offs), GAL_OP_FREE, dtype=np.uint32), offs, seed=self.seed)
        self.live.clear()
        return r
