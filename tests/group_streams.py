"""Random multi-lane op streams for gal_group_replay / the oracle's group replay.

An op is (kind, slot, lanes, args, tids): args and tids are 64-entry lists
indexed by lane. A malloc op's sizes are all on the slice path or all above
it; a free op of several lanes frees slices or whole blocks only (a segment
run is freed by a one-lane op). The stream is built against a live oracle so
that every free names a malloc that succeeded.
"""
import random

from oracle import oracle

NULL = (1 << 64) - 1


def lane_mask(rng):
    kind = rng.random()
    if kind < 0.15:
        return (1 << 64) - 1
    if kind < 0.3:
        return 1 << rng.randrange(64)
    if kind < 0.45:
        lo = rng.randrange(64)
        hi = rng.randrange(lo, 64)
        return ((1 << (hi - lo + 1)) - 1) << lo
    density = rng.choice((0.1, 0.3, 0.5, 0.8))
    m = 0
    for lane in range(64):
        if rng.random() < density:
            m |= 1 << lane
    return m or 1 << rng.randrange(64)


def build(o, seed, n_ops, slice_sizes, large_sizes=(), free_p=0.4, large_p=0.05, slots=256, seg=16 << 20,
          mn=16, mx=4096, spread_p=0.0):
    """Generate n_ops ops, running each on the oracle `o` as it is made.
    Returns (ops, results) with results an (n_ops, 64) list of offsets.
    A malloc op of slice sizes is GAL_OP_MALLOC_SPREAD (kind 2, the claims
    malloc makes under load) with probability spread_p."""
    rng = random.Random(seed)
    large_sizes = [x for x in large_sizes if oracle.route(seg, mn, mx, x)[0] != 0]
    ops, results = [], []
    live = []  # (flat index, offset, is_segment_run)
    for i in range(n_ops):
        lanes = lane_mask(rng)
        if live and rng.random() < free_p:
            lanes_list = [lane for lane in range(64) if (lanes >> lane) & 1]
            rng.shuffle(live)
            picks = []
            for entry in live:
                if len(picks) == len(lanes_list):
                    break
                if entry[2] and (len(lanes_list) > 1 or picks):
                    continue
                picks.append(entry)
                if entry[2]:
                    break
            lanes_list = lanes_list[: len(picks)]
            if not picks:
                continue
            for entry in picks:
                live.remove(entry)
            args, offs, lanes = [0] * 64, [0] * 64, 0
            for lane, entry in zip(lanes_list, picks):
                args[lane] = entry[0]
                offs[lane] = entry[1]
                lanes |= 1 << lane
            assert o.group_free(lanes, offs) == 0
            ops.append((1, 0, lanes, args, [0] * 64))
            results.append([NULL if not (lanes >> k) & 1 else 0 for k in range(64)])
            continue
        slot = rng.randrange(slots)
        tids = [rng.randrange(1 << 24) for _ in range(64)]
        if large_sizes and rng.random() < large_p:
            # a few lanes asking for blocks or segments
            m = 0
            for _ in range(rng.randint(1, 3)):
                m |= 1 << rng.randrange(64)
            lanes = m
            sizes = [rng.choice(large_sizes) for _ in range(64)]
        elif rng.random() < 0.3:
            sizes = [rng.choice(slice_sizes)] * 64
        else:
            sizes = [rng.choice(slice_sizes) for _ in range(64)]
        spread = oracle.route(seg, mn, mx, sizes[(lanes & -lanes).bit_length() - 1])[0] == 0 and rng.random() < spread_p
        out = o.group_malloc(lanes, sizes, tids, slot, spread=spread)
        idx = len(ops)
        ops.append((2 if spread else 0, slot, lanes, sizes, tids))
        row = [NULL] * 64
        for lane in range(64):
            if (lanes >> lane) & 1:
                row[lane] = out[lane]
                if out[lane] != NULL:
                    path = oracle.route(seg, mn, mx, sizes[lane])[0]
                    live.append((idx * 64 + lane, out[lane], path == 2))
        results.append(row)
    return ops, results, live
