"""Allocator invariants, shared by the oracle tests and the GPU tests.

Live-set checks (any time): no two live allocations overlap, each starts on
its granted size's alignment, each lies inside the pool.
Quiescent checks (no allocation live, no call in flight) on a snapshot:
  - every item of a class tree is reachable from its summaries
  - a segment is free (chunk id 0xFFFF) exactly when its segment-tree bit is set
  - a class tree only lists segments formatted for that class
  - per formatted segment: active == nb - 1 - queue_alloc + queue_free,
    queue_final == queue_free, and the blocks granted and not returned are
    exactly the pinned blocks that live in it
  - a pinned block has returned every slice it handed out
    (free_counter == slices claimed) and carries its class in malloc_counter;
    every other block has free_counter == 0
  - no lock bit is held
"""
import numpy as np

from gallatin_amd import snapshot as snap

NULL = (1 << 64) - 1


def granted_bytes(seg, mn, mx, size, route):
    path, tree, count = route(seg, mn, mx, size)
    if path == 0:
        return (mn << tree) * count, (mn << tree) if count == 1 else mx
    if path == 1:
        b = (mn << tree) * 4096
        return b, b
    return count * seg, seg


def check_live(offsets, sizes, pool_bytes, seg, mn, mx, route):
    """offsets/sizes: numpy arrays of live allocations (byte offsets)."""
    offsets = np.asarray(offsets, dtype=np.uint64)
    sizes = np.asarray(sizes, dtype=np.uint64)
    if len(offsets) == 0:
        return
    uniq, inv = np.unique(sizes, return_inverse=True)
    table = np.array([granted_bytes(seg, mn, mx, int(s), route) for s in uniq.tolist()], dtype=np.uint64)
    granted, align = table[inv, 0], table[inv, 1]
    assert np.all(offsets % align == 0), "misaligned allocation"
    assert np.all(offsets + granted <= np.uint64(pool_bytes)), "allocation outside the pool"
    order = np.argsort(offsets, kind="stable")
    o, g = offsets[order], granted[order]
    overlap = np.nonzero(o[:-1] + g[:-1] > o[1:])[0]
    if len(overlap):
        sz = sizes[order]
        pairs = [
            (int(o[k]), int(sz[k]), int(o[k + 1]), int(sz[k + 1]), int(order[k]), int(order[k + 1]))
            for k in overlap[:8]
        ]
        segs = sorted({int(o[k]) // seg for k in overlap})
        raise AssertionError(
            f"{len(overlap)} overlapping allocations in segments {segs[:16]}; first pairs "
            f"(off, size, off, size, idx, idx): {pairs}"
        )


def pinned_slots_of(c):
    """pinned slots of class c: 128 >> c, at least 4 (reference shared_block_storage.cuh:251-274)"""
    return max(4, 128 >> c)


def empty_pinned_counts(image):
    """Per class, the pinned slots holding 0 in a snapshot (the slot region of
    a class is its pinned slots followed by overflow slots up to 64)."""
    s = snap.parse(image)
    slots = s["slots"].astype(np.int64)
    out, base = [], 0
    for c in range(s["num_trees"]):
        n = pinned_slots_of(c)
        out.append(int(np.count_nonzero(slots[base:base + n] == 0)))
        base += max(n, 64)
    return out


def bits_of(words):
    out = []
    for w, val in enumerate(np.asarray(words, dtype=np.uint64).tolist()):
        while val:
            low = val & -val
            out.append(w * 64 + low.bit_length() - 1)
            val ^= low
    return out


def check_quiescent(image, seg_bytes, min_size):
    s = snap.parse(image)
    nseg, bps, T = s["num_segments"], s["blocks_per_segment"], s["num_trees"]
    chunk = s["chunk_ids"].astype(np.int64)
    assert s["locks"] == 0, "a lock bit is held"

    free_bits = np.zeros(nseg, dtype=bool)
    free_bits[bits_of(s["segment_tree"][-1])] = True
    assert np.array_equal(free_bits, chunk == 0xFFFF), "segment tree and chunk ids disagree"

    for c in range(T):
        layers = s["sub_trees"][c]
        listed = bits_of(layers[-1])
        assert all(chunk[x] == c for x in listed), f"class {c} tree lists a segment of another class"
        # every listed segment is reachable: a word that holds items has its
        # summary bit set (class trees only change by single inserts/removes)
        for level, (upper, lower) in enumerate(zip(layers[:-1], layers[1:])):
            summary = set(bits_of(upper))
            lower_vals = np.asarray(lower, dtype=np.uint64).tolist()
            holding = [w for w, val in enumerate(lower_vals) if val]
            lost = [w for w in holding if w not in summary]
            assert not lost, (f"class {c} tree: words {lost[:8]} of layer {level + 1} hold items "
                              f"({[hex(lower_vals[w]) for w in lost[:4]]}) but their summary bits in layer "
                              f"{level} are clear (layer {level}: {[hex(int(x)) for x in np.asarray(upper)[:8]]})")

    # blocks held by slots: each slot's entry and the block fetched ahead for it
    slots = s["slots"].astype(np.int64)
    ahead = s["ahead"].astype(np.int64)
    assert not np.any(slots == 0xFFFFFFFF), "a slot is still busy"
    assert not np.any(ahead == 0xFFFFFFFF), "a fetch ahead is unfinished"
    pinned = np.concatenate([slots[slots > 0], ahead[ahead > 0]]) - 1
    assert len(np.unique(pinned)) == len(pinned), "a block is pinned twice"
    pinned_per_seg = np.bincount(pinned // bps, minlength=nseg)

    active = s["active"].astype(np.int64)
    qa, qf, qfin = (s[k].astype(np.int64) for k in ("queue_alloc", "queue_free", "queue_final"))
    for x in range(nseg):
        c = chunk[x]
        if c < T:
            nb = seg_bytes // ((min_size << c) * 4096)
            assert active[x] == nb - 1 - qa[x] + qf[x], f"segment {x}: active count off"
            assert qfin[x] == qf[x], f"segment {x}: unfinished queue write"
            outstanding = nb - 1 - active[x]
            if outstanding != pinned_per_seg[x]:
                lo = x * bps
                blk = [(i, int(s["malloc_counter"][lo + i]), int(s["free_counter"][lo + i])) for i in range(nb)]
                raise AssertionError(
                    f"segment {x} (class {c}, nb {nb}): {outstanding} blocks out, {pinned_per_seg[x]} pinned; "
                    f"active {active[x]} q_alloc {qa[x]} q_free {qf[x]} q_final {qfin[x]}; "
                    f"queue {s['queues'][lo:lo + nb].tolist()}; pinned here "
                    f"{sorted(int(b - lo) for b in pinned if b // bps == x)}; blocks (idx, malloc, free) {blk}"
                )
        elif c == 0xFFFF:
            assert active[x] == -1, f"free segment {x} has active {active[x]}"
            assert pinned_per_seg[x] == 0

    mc, fc = s["malloc_counter"].astype(np.int64), s["free_counter"].astype(np.int64)
    is_pinned = np.zeros(nseg * bps, dtype=bool)
    is_pinned[pinned] = True
    count = mc & ((1 << 20) - 1)
    assert np.all(fc[~is_pinned] == 0), "a block that is not pinned still counts frees"
    assert np.all(fc[is_pinned] == count[is_pinned]), "a pinned block has slices out"
    seg_of = pinned // bps
    assert np.all((mc[pinned] >> 20) == chunk[seg_of]), "pinned block of the wrong class"
    base = 0
    for c in range(T):
        region = ahead[base:base + max(pinned_slots_of(c), 64)]
        base += max(pinned_slots_of(c), 64)
        fetched = region[region > 0] - 1
        assert np.all(count[fetched] == 0), f"a block fetched ahead for class {c} has slices claimed"
        assert np.all((mc[fetched] >> 20) == c), f"a block of another class fetched ahead for class {c}"
    return s
