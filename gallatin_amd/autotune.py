"""Autotuner for the allocator's configuration on a given workload.

What it chooses:
  * the instantiation (segment_size, min_size, max_size), among the triples
    compiled into the library that are legal for the workload: powers of
    two, a segment holds at least one block of the top class
    (segment_size >= 4096 * max_size, reference gallatin.cuh:300), at most
    30 classes, min_size fixed by the caller (128 for slab-sized records);
  * for batched service, the claim-group width (lanes that claim together)
    and the run length (requests a lane claims with one bulk claim).

How: every candidate triple gets a fresh allocator over the trial pool; the
trial is the workload itself at reduced size (a slab hash build with
`slab_keys` keys, then `replays` batched services of `batch_requests` mixed
requests for each (width, run) tried). The search has two stages: every
triple is measured with one probe setting (the last width and run given),
then the `finalists` best triples are measured over the whole (width, run)
grid. The highest combined rate (slab inserts plus batched mallocs+frees,
over the summed kernel time) wins. A `measure` function can stand in for the
GPU trial (tests). Results can be cached in a JSON file keyed by the
workload and the candidate set.
"""
import json
import os

from . import FORM_GRAPH, Batch, Gallatin, SlabTable, supported_triples

WIDTHS = (8, 16, 32, 64)
RUNS = (1, 4, 16)


def legal(seg, mn, mx):
    p2 = all(v > 0 and v & (v - 1) == 0 for v in (seg, mn, mx))
    classes = mx.bit_length() - mn.bit_length() + 1
    return p2 and mn <= mx and seg >= 4096 * mx and classes <= 30


def candidates(min_size, segment_sizes=None):
    out = []
    for seg, mn, mx in supported_triples():
        if mn != min_size or not legal(seg, mn, mx):
            continue
        if segment_sizes is not None and seg not in segment_sizes:
            continue
        out.append((seg, mn, mx))
    return sorted(out)


def gpu_trial(triple, pool_bytes, slab_keys, slab_buckets, batch_requests, size_classes, replays, widths, runs,
              seed=1):
    """Measure one triple on the GPU: slab build, then every (width, run)."""
    seg, mn, mx = triple
    g = Gallatin.generate_on_device(pool_bytes, seed, segment_size=seg, min_size=mn, max_size=mx)
    try:
        slab = {"keys": 0, "kernel_ms": 0.0}
        table = None
        if slab_keys:
            table = SlabTable(g, slab_buckets, seed=seed)
            slab = table.insert(slab_keys, seed)
            if slab["failed"] or slab["alloc_errors"]:
                raise RuntimeError(f"slab build failed on {triple}: {slab}")
        batch = {}
        if batch_requests:
            with Batch(g, batch_requests) as b:
                b.generate(batch_requests, seed=seed, size_classes=size_classes)
                for w in widths:
                    for r in runs:
                        res = b.run(FORM_GRAPH, width=w, run=r, replays=replays)
                        if res["misses"] or res["tag_errors"] or res["alloc_errors"]:
                            continue
                        batch[(w, r)] = {"ops": res["mallocs"] + res["frees"], "ms": res["total_ms"]}
        if table is not None:
            table.destroy()
        return {"slab": {"ops": slab["keys"], "ms": slab["kernel_ms"]}, "batch": batch}
    finally:
        g.free_on_device()


def tune(pool_bytes, min_size=128, segment_sizes=(4 << 20, 16 << 20, 64 << 20), slab_keys=1 << 22,
         slab_buckets=1 << 16, batch_requests=1 << 22, size_classes=13, replays=3, widths=WIDTHS, runs=RUNS,
         measure=None, cache=None, finalists=2):
    """Returns {"triple", "width", "run", "rate", "trials"}."""
    cands = candidates(min_size, segment_sizes)
    if not cands:
        raise ValueError(f"no legal instantiation with min_size={min_size} and segment sizes {segment_sizes}")
    key = json.dumps([pool_bytes, min_size, cands, slab_keys, slab_buckets, batch_requests, size_classes, replays,
                      list(widths), list(runs), finalists])
    if cache and os.path.exists(cache):
        with open(cache) as f:
            stored = json.load(f)
        if key in stored:
            hit = stored[key]
            hit["triple"] = tuple(hit["triple"])
            return hit
    measure = measure or gpu_trial

    def score(triple, ws, rs):
        m = measure(triple, pool_bytes, slab_keys, slab_buckets, batch_requests, size_classes, replays, ws, rs)
        if not m["batch"] and batch_requests:
            return {"triple": triple, "rate": 0.0, "note": "every batch trial missed"}
        if m["batch"]:
            (w, r), b = max(m["batch"].items(), key=lambda kv: kv[1]["ops"] / kv[1]["ms"])
        else:
            (w, r), b = (ws[-1], rs[-1]), {"ops": 0, "ms": 0.0}
        ops = m["slab"]["ops"] + b["ops"]
        ms = m["slab"]["ms"] + b["ms"]
        return {"triple": triple, "width": w, "run": r, "rate": ops / (ms * 1e-3) if ms > 0 else 0.0,
                "slab_rate": m["slab"]["ops"] / (m["slab"]["ms"] * 1e-3) if m["slab"]["ms"] > 0 else None,
                "batch_rate": b["ops"] / (b["ms"] * 1e-3) if b["ms"] > 0 else None}

    probe = [score(t, (widths[-1],), (runs[-1],)) for t in cands]
    trials = list(probe)
    ranked = sorted((p for p in probe if p["rate"] > 0), key=lambda p: -p["rate"])
    best = None
    for p in ranked[:finalists]:
        full = score(p["triple"], widths, runs) if (len(widths), len(runs)) != (1, 1) else p
        full["stage"] = "grid"
        trials.append(full)
        if best is None or full["rate"] > best["rate"]:
            best = full
    if best is None:
        raise RuntimeError("no candidate served the trial without misses")
    out = dict(best)
    out["trials"] = trials
    if cache:
        stored = {}
        if os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
        stored[key] = out
        with open(cache, "w") as f:
            json.dump(stored, f)
    return out
