"""Parse a metadata snapshot image (layout documented in include/gallatin_amd.h)."""
import numpy as np


def parse(buf, num_trees=None):
    buf = np.asarray(buf, dtype=np.uint8)
    head = buf[:64].view(np.uint64)
    nseg, bps, trees, layers, slots, locks = (int(x) for x in head[:6])
    off = 64

    # words per layer, top (0) to items (layers-1)
    words = []
    u = nseg
    for _ in range(layers):
        words.append((u - 1) // 64 + 1)
        u = (u - 1) // 64 + 1
    words = words[::-1]

    def take(nbytes, dtype):
        nonlocal off
        arr = buf[off : off + nbytes].view(dtype)
        off += nbytes
        return arr

    def tree():
        return [take(w * 8, np.uint64) for w in words]

    seg_tree = tree()
    sub_trees = [tree() for _ in range(trees)]
    chunk = take(nseg * 2, np.uint16)
    off += ((nseg * 2 + 7) // 8) * 8 - nseg * 2
    nblocks = nseg * bps
    blocks = take(nblocks * 8, np.uint32).reshape(nblocks, 2)
    queues = take(nblocks * 4, np.uint32)
    segs = take(nseg * 16, np.uint32).reshape(nseg, 4)
    slot_arr = take(slots * 4, np.uint32)
    ahead_arr = take(slots * 4, np.uint32)
    return {
        "num_segments": nseg,
        "blocks_per_segment": bps,
        "num_trees": trees,
        "layers": layers,
        "locks": locks,
        "segment_tree": seg_tree,
        "sub_trees": sub_trees,
        "chunk_ids": chunk,
        "malloc_counter": blocks[:, 0],
        "free_counter": blocks[:, 1],
        "queues": queues,
        "active": segs[:, 0].view(np.int32),
        "queue_alloc": segs[:, 1],
        "queue_free": segs[:, 2],
        "queue_final": segs[:, 3],
        "slots": slot_arr,
        "ahead": ahead_arr,
    }


def popcount64(words):
    w = np.asarray(words, dtype=np.uint64)
    return int(sum(bin(int(x)).count("1") for x in w))


def first_difference(a, b):
    """Name and index of the first field where two parsed snapshots differ."""
    pa, pb = parse(a), parse(b)
    for key in pa:
        va, vb = pa[key], pb[key]
        if isinstance(va, list):
            for i, (x, y) in enumerate(zip(va, vb)):
                if isinstance(x, list):
                    for j, (xx, yy) in enumerate(zip(x, y)):
                        d = np.nonzero(np.asarray(xx) != np.asarray(yy))[0]
                        if len(d):
                            return f"{key}[{i}][layer {j}] word {d[0]}: {xx[d[0]]:#x} vs {yy[d[0]]:#x}"
                else:
                    d = np.nonzero(np.asarray(x) != np.asarray(y))[0]
                    if len(d):
                        return f"{key}[layer {i}] word {d[0]}: {x[d[0]]:#x} vs {y[d[0]]:#x}"
        elif isinstance(va, np.ndarray):
            d = np.nonzero(va != vb)[0]
            if len(d):
                return f"{key}[{d[0]}]: {va[d[0]]} vs {vb[d[0]]}"
        elif va != vb:
            return f"{key}: {va} vs {vb}"
    return None
