#!/usr/bin/env python3
"""Benchmark: device mallocs+frees/sec on one MI355X (mixed sizes), p99 alloc latency.

Workload (one step): Gallatin<16 MiB, 16 B, 4096 B> over an 8 GiB pool;
10 rounds of 2^22 threads, each thread mallocs a power-of-two size 16 B..4 KiB
drawn log-uniform from (seed, round, thread), tags it; after every round each
live pointer is freed with probability 1/2 (seeded coin, tag checked first);
after the last round everything still live is freed. Inputs are generated on
the device from the seed, so nothing is transferred inside the timed region.

Prints ONE JSON line. Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import glob
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "device mallocs+frees/sec on 1 MI355X (mixed sizes); p99 alloc latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per malloc+free pair on the slice path, one request per
# lane (no group sharing), from the metadata layout
# (gallatin_amd/include/gallatin/allocators/memory_table.hpp, gallatin.hpp):
SLOT_ENTRY = 4       # u32 pinned-slot entry, read by the claim
MALLOC_COUNTER = 8   # u64 Block::malloc_counter, read and written by the claim's atomic add
FREE_COUNTER = 4     # u32 Block::free_counter, read and written by the free's atomic add
CHUNK_ID = 4         # u32 chunk id of the segment, read by free
TAG = 8              # u64 payload tag, written after malloc and read before free
BYTES_MALLOC_SIDE = SLOT_ENTRY + 2 * MALLOC_COUNTER + TAG  # 28 B per successful malloc
BYTES_FREE_SIDE = CHUNK_ID + 2 * FREE_COUNTER + TAG  # 20 B per free
# Counter figures of the default workload's kernels, per launch averaged over
# one step, measured by rocprofv3 on the kernel sources whose digest is
# `sources_sha` (the digest of KERNEL_SOURCES below; the figures are reported
# only while the sources still have that digest):
#   traffic: FETCH_SIZE and WRITE_SIZE runs (KiB per dispatch) scaled by what
#     the counters read for a 1 GiB device copy (fetch x2.0, write x1.0),
#     profiles/mixed_slices_8GiB_hbm_traffic.md;
#   atomic_requests: TCC_ATOMIC_sum (every agent-scope atomic is performed
#     beyond the L2: TCC_EA0_ATOMIC_sum reads the same; one request serves
#     the lanes of one instruction within one 64-B span),
#     profiles/mixed_slices_8GiB_tcc_atomics.md.
MEASURED = {
    "sources_sha": "4ea174eecb2597b5",
    "commit": "1907a23",
    "traffic": {"mixed_malloc_kernel": 207444995, "mixed_free_kernel": 205178647},
    "atomic_requests": {"mixed_malloc_kernel": 335218, "mixed_free_kernel": 589108},
}
# the sources the mixed workload's kernels are compiled from: their kernels
# (workloads.hip) and the device allocator (headers)
KERNEL_SOURCES = ("gallatin_amd/csrc/workloads.hip", "gallatin_amd/include/gallatin/*.hpp",
                  "gallatin_amd/include/gallatin/allocators/*.hpp")
DEFAULT_WORKLOAD = (8, 22, 10)  # pool GiB, log2 threads per round, rounds


def sources_sha():
    """Digest of the kernel sources the library is built from."""
    h = hashlib.sha256()
    for pattern in KERNEL_SOURCES:
        for path in sorted(glob.glob(os.path.join(ROOT, pattern))):
            h.update(os.path.relpath(path, ROOT).encode())
            with open(path, "rb") as f:
                h.update(f.read())
    return h.hexdigest()[:16]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(pool, seed, threads, rounds):
    from oracle import oracle

    ops, miss, sec = oracle.bench_mixed(pool, seed, threads, rounds)
    sysm = oracle.bench_sysmalloc(1 << 22, 0, seed)
    return {
        "value": ops / sec if sec > 0 else None,
        "unit": "mallocs+frees/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle (sequential C++ restatement) on the same mixed workload at {threads} threads/round x "
        f"{rounds} rounds over {pool >> 30} GiB: {ops} successful mallocs+frees, {miss} misses, {sec:.2f} s",
        "system_malloc": {
            "value": sysm["ops"] / sysm["seconds"] if sysm["seconds"] > 0 else None,
            "unit": "mallocs+frees/s",
            "cores": sysm["threads"],
            "p50_ns": sysm["p50_ns"],
            "p99_ns": sysm["p99_ns"],
            "sample": f"glibc malloc/free, {1 << 22} ops per thread, log-uniform 16 B-4 KiB, tag write+check",
        },
    }


def rate(ops, ms):
    return ops / (ms * 1e-3) if ms > 0 else None


def plain(seed, launches, cpu):
    """The reference's global test kernel at full size: a 1 GiB global
    allocator, one kernel of 2^20 threads, each global_malloc(16 B), writes a
    tag drawn from (seed, thread), checks it, global_free."""
    import gallatin_amd as ga

    n = 1 << 20
    ga.init_global_allocator(1 << 30, seed, print_info=False)
    try:
        ga.global_plain(n, 16, seed, ga.GAL_VERIFY_STORE)  # warm-up
        rs = [ga.global_plain(n, 16, seed + 1 + k, ga.GAL_VERIFY_STORE) for k in range(launches)]
        dev, _ = ga.global_handles()
        errors = dev.errors()
    finally:
        ga.free_global_allocator()
    ops = sum(r["mallocs"] + r["frees"] for r in rs)
    ms = sum(r["total_ms"] for r in rs)
    out = {
        "workload": "1 GiB global allocator (Gallatin<16MiB,16,4096>), one kernel of 2^20 threads: "
        "global_malloc(16 B), tag write, check, global_free",
        "launches": launches,
        "mallocs_frees_per_s": rate(ops, ms),
        "kernel_avg_us": ms / launches * 1e3,
        "misses": sum(r["misses"] for r in rs),
        "tag_errors": sum(r["tag_errors"] for r in rs),
        "alloc_errors": sum(r["alloc_errors"] for r in rs),
        "errors": errors,
    }
    if cpu:
        from oracle import oracle

        sysm = oracle.bench_sysmalloc(1 << 22, 0, seed, size=16)
        out["cpu_baseline"] = {
            "value": sysm["ops"] / sysm["seconds"] if sysm["seconds"] > 0 else None,
            "unit": "mallocs+frees/s",
            "cores": sysm["threads"],
            "kind": "system",
            "sample": f"glibc malloc(16)/tag write/check/free, {1 << 22} per thread on every host thread",
            "p50_ns": sysm["p50_ns"],
            "p99_ns": sysm["p99_ns"],
        }
    log(f"[bench] plain: {out}")
    return out


def forms(seed):
    """The forms of service on one request mix (16 B-4 KiB, log-uniform powers
    of two): one thread per request (malloc, tag, check, free), the binned
    persistent launch, the same launch replayed as a hipGraph, and the
    serving loop over sequences of pages of those sizes."""
    import gallatin_amd as ga

    n = 1 << 22
    out = {"requests": n, "sizes": "16 B-4 KiB, 16 << k, k uniform in [0, 9)"}
    g = ga.Gallatin.generate_on_device(8 << 30, seed)
    try:
        with ga.Batch(g, n) as b:
            b.generate(n, seed=seed, size_classes=9)
            for name, form, width, run in (("per_thread", ga.FORM_PER_THREAD, 64, 1),
                                           ("batched", ga.FORM_BATCHED, 64, 32),
                                           ("graph", ga.FORM_GRAPH, 64, 32)):
                b.run(form, width=width, run=run, replays=1)  # warm-up
                r = b.run(form, width=width, run=run, replays=5, verify=ga.GAL_VERIFY_STORE)
                out[name] = {"mallocs_frees_per_s": rate(r["mallocs"] + r["frees"], r["total_ms"]),
                             "misses": r["misses"], "tag_errors": r["tag_errors"], "alloc_errors": r["alloc_errors"],
                             "width": width, "run": run}
                log(f"[bench] form {name}: {out[name]}")
        out["errors"] = g.errors()
    finally:
        g.free_on_device()
    # Serving: sequences of pages of the same sizes, all arrived at once, on
    # the serving workload's 200 GiB pool. The sustained rate counts pages
    # written and freed in a window that ends before any slot finds the
    # queues empty (the loop is saturated for all of it); the whole-run rate
    # also carries the drain of the last, longest sequences.
    g = ga.Gallatin.generate_on_device(200 << 30, seed)
    try:
        for name, med, cap, n, slots, window in (("serving", 1024, 32768, 1 << 19, 65536, 0.1),
                                                  ("serving_short", 16, 1024, 1 << 23, 65536, 0.04)):
            reqs = ga.generate_requests(n, 0.0, median_pages=med, sigma=1.0, max_pages=cap, page_bytes=0,
                                        size_classes=9, seed=seed)
            with ga.Serving(g, len(reqs)) as s:
                s.load(reqs)
                r = s.run(slots=slots, run=64, deadline_s=60.0, window_s=window)
            saturated = r["first_dry_ms"] < 0 or r["first_dry_ms"] > r["window_ms"]
            out[name] = {
                "mallocs_frees_per_s": r["window_pages"] / (r["window_ms"] * 1e-3) if saturated else None,
                "whole_run_mallocs_frees_per_s": rate(r["page_mallocs"] + r["page_frees"], r["kernel_ms"]),
                "window_ms": r["window_ms"], "first_dry_ms": r["first_dry_ms"],
                "workload": f"{n} sequences all arrived at once, lognormal length (median {med}, sigma 1, cap {cap} "
                f"pages), page sizes as above, {slots} slots, runs of 64, 200 GiB pool",
                "page_malloc_p50_ns": r["page_malloc_p50_ns"], "page_malloc_p99_ns": r["page_malloc_p99_ns"],
                "done": r["requests_done"], "aborted": r["requests_aborted"], "alloc_errors": r["alloc_errors"]}
            log(f"[bench] form {name}: {out[name]}")
        out["serving_errors"] = g.errors()
    finally:
        g.free_on_device()
    base = out["per_thread"]["mallocs_frees_per_s"]
    out["vs_per_thread"] = {k: (out[k]["mallocs_frees_per_s"] / base if out[k]["mallocs_frees_per_s"] else None)
                            for k in ("batched", "graph", "serving", "serving_short")}
    # the serving forms over their whole run (ramp-down at the end included)
    out["vs_per_thread_whole_run"] = {k: out[k]["whole_run_mallocs_frees_per_s"] / base
                                      for k in ("serving", "serving_short")}
    return out


def serving(seed, requests, rate_per_s, run=64, sample=20000, grace_s=60.0):
    """Sequences arriving by a Poisson process, growing by 16 KiB pages, admitted
    by free pool bytes, over a 200 GiB pool. `run` pages are claimed per
    allocator call (1: one malloc per page). The checksums of a seeded sample
    of requests are checked against a host computation."""
    import numpy as np

    import gallatin_amd as ga
    from tests import serving_ref

    g = ga.Gallatin.generate_on_device(200 << 30, seed)
    try:
        reqs = ga.generate_requests(requests, rate_per_s, median_pages=1024, sigma=1.0, max_pages=32768,
                                    page_bytes=16384, seed=seed)
        with ga.Serving(g, len(reqs)) as s:
            s.load(reqs)
            r = s.run(slots=16384, run=run, deadline_s=float(reqs["arrival_ns"][-1]) * 1e-9 + grace_s)
            sums, _, _ = s.results()
        pick = np.random.default_rng(seed).choice(len(reqs), size=min(sample, len(reqs)), replace=False)
        done = pick[sums[pick] != 0]
        expect = serving_ref.checksums_c(reqs["id"][done], reqs["pages"][done], reqs["page_bytes"][done])
        sec = r["kernel_ms"] * 1e-3
        out = {
            "workload": f"200 GiB pool, {requests} requests, Poisson {rate_per_s:.0f}/s, lognormal length (median "
            f"1024, sigma 1, cap 32768) of 16 KiB pages, admission by free pool bytes, 16384 slots, runs of {run}",
            "run": run,
            "checksum_sample": int(len(pick)),
            "checksum_checked": int(len(done)),
            "checksum_mismatches": int(np.count_nonzero(sums[done] != expect)),
            "arrival_span_s": float(reqs["arrival_ns"][-1]) * 1e-9,
            "requests_per_s": r["requests_done"] / sec if sec > 0 else None,
            "page_mallocs_frees_per_s": rate(r["page_mallocs"] + r["page_frees"], r["kernel_ms"]),
            "p50_page_malloc_ns": r["page_malloc_p50_ns"],
            "p99_page_malloc_ns": r["page_malloc_p99_ns"],
        }
        out.update({k: r[k] for k in ("requests_done", "requests_aborted", "page_mallocs", "page_retries",
                                      "admission_refusals", "peak_active", "peak_committed", "queue_p50_ns",
                                      "queue_p99_ns", "latency_p50_ns", "latency_p99_ns", "alloc_errors",
                                      "kernel_ms")})
        out["errors"] = g.errors()
        log(f"[bench] serving: {out}")
        return out
    finally:
        g.free_on_device()


def blocks(seed):
    """Block and multi-segment churn with the fragmentation figure per round."""
    import gallatin_amd as ga
    from gallatin_amd.workloads import BlockChurn

    g = ga.Gallatin.generate_on_device(64 << 30, seed)
    try:
        w = BlockChurn(g, requests=1 << 16, rounds=20, seed=seed)
        hist = w.run()
        d = w.drain()
        ops = sum(h["mallocs"] + h["frees"] for h in hist) + (d["frees"] if d else 0)
        ms = sum(h["kernel_ms"] for h in hist) + (d["malloc_kernel_ms"] if d else 0.0)
        out = {
            "workload": "64 GiB pool, 2^16 requests of floor(2^U(16,28)) bytes over 20 rounds, 60% mallocs / "
            "40% frees per round in one concurrent kernel",
            "mallocs_frees_per_s": rate(ops, ms),
            "mallocs": sum(h["mallocs"] for h in hist),
            "misses": sum(h["misses"] for h in hist),
            "tag_errors": sum(h["tag_errors"] for h in hist),
            "largest_request_bytes_per_round": [h["largest_request_bytes"] for h in hist],
            "errors": g.errors(),
        }
        log(f"[bench] blocks: {out}")
        return out
    finally:
        g.free_on_device()


def slab(seed):
    """Slab hash build on an autotuned allocator, then hipGraph-replayed batches."""
    import gallatin_amd as ga
    from gallatin_amd import autotune

    pool = 32 << 30
    best = autotune.tune(pool, slab_keys=1 << 22, slab_buckets=1 << 16, batch_requests=1 << 22, size_classes=13,
                         replays=2, widths=(16, 64), runs=(16, 32))
    seg, mn, mx = best["triple"]
    g = ga.Gallatin.generate_on_device(pool, seed, segment_size=seg, min_size=mn, max_size=mx)
    try:
        t = ga.SlabTable(g, 1 << 20, seed=seed)
        ins = t.insert(1 << 26, key_seed=seed)
        q = t.query(1 << 26, key_seed=seed)
        n = 1 << 24
        with ga.Batch(g, n) as b:
            b.generate(n, seed=seed, size_classes=13)
            r = b.run(ga.FORM_GRAPH, width=best["width"], run=best["run"], replays=100)
        out = {
            "workload": "32 GiB pool; 2^26 random u64 keys into a 2^20-bucket slab hash (120-byte slabs); then "
            "2^24 requests of 16 B-64 KiB per hipGraph replay, 100 replays, on the live table",
            "autotuned": {"segment_size": seg, "min_size": mn, "max_size": mx, "width": best["width"],
                          "run": best["run"],
                          "trials": [{"triple": list(x["triple"]), "width": x.get("width"), "run": x.get("run"),
                                      "rate": x["rate"], "stage": x.get("stage", "probe")} for x in best["trials"]]},
            "slab_inserts_per_s": rate(ins["keys"], ins["kernel_ms"]),
            "slabs": ins["slabs"],
            "max_chain": ins["max_chain"],
            "insert_failed": ins["failed"],
            "query_missed": q["failed"] + q["wrong"],
            "batch_mallocs_frees_per_s": rate(r["mallocs"] + r["frees"], r["total_ms"]),
            "batch_misses": r["misses"],
            "batch_tag_errors": r["tag_errors"],
        }
        t.destroy()
        out["errors"] = g.errors()
        log(f"[bench] slab: {out}")
        return out
    finally:
        g.free_on_device()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pool-gib", type=int, default=8)
    ap.add_argument("--threads-log2", type=int, default=22)
    ap.add_argument("--rounds", type=int, default=10)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="skip the other workloads (forms, serving, blocks, slab)")
    ap.add_argument("--serving-requests", type=int, default=1000000, help="requests of the serving workload")
    ap.add_argument("--serving-rate", type=float, default=50000.0)
    ap.add_argument("--cpu-threads-log2", type=int, default=22,
                    help="oracle sample size (threads per round; the workload's own 2^22 is about 40 s of one core)")
    ap.add_argument(
        "--copy-calibration",
        type=int,
        default=0,
        metavar="MIB",
        help="only run 3 plain device copies of MIB MiB (calibrates FETCH_SIZE/WRITE_SIZE counters) and exit",
    )
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(
            "bench.py measures one GPU: the allocator is per-device state with no cross-GPU traffic; "
            "run with --gpus 1",
            file=sys.stderr,
        )
        sys.exit(2)

    import gallatin_amd as ga

    pool = args.pool_gib << 30
    threads = 1 << args.threads_log2
    rounds = args.rounds
    log(f"[bench] creating Gallatin<16MiB,16,4096> pool={pool} threads/round={threads} rounds={rounds}")
    g = ga.Gallatin.generate_on_device(pool, args.seed)

    def step(seed, tag, rare_paths=False):
        """One step, round by round for the log: reset, every round (malloc
        kernel + free kernel, then a device read of the round's counters),
        drain. rare_paths: also read the allocator's rare-path counters after
        each round (one more device read per round). Used for the warm-up."""
        g.mixed_reset(threads, rounds, seed)
        total = None
        before = g.stats() if rare_paths else None
        for r in range(rounds):
            res = g.mixed_round(r)
            line = (f"[bench] {tag} round {r}: mallocs {res['mallocs']} frees {res['frees']} misses {res['misses']} "
                    f"malloc_ms {res['malloc_kernel_ms']:.2f} free_ms {res['free_kernel_ms']:.2f}")
            if rare_paths:
                now = g.stats()
                line += f" rare paths { {k: now[k] - before[k] for k in now if now[k] != before[k]} }"
                before = now
            log(line)
            total = res if total is None else {k: total[k] + res[k] for k in total}
        res = g.mixed_drain()
        total = {k: total[k] + res[k] for k in total}
        return total

    for w in range(args.warmup):
        r = step(args.seed + 1000 + w, f"warmup {w}", rare_paths=True)
        log(f"[bench] warmup {w}: {r}")

    results = []
    # the latency samples of every timed step, not only the last one's
    g.latency_keep(args.steps)
    t0 = time.perf_counter()
    for k in range(args.steps):
        # the same launches in one call: every round's malloc and free kernel
        # and the drain back to back on the library's stream, one
        # synchronisation at the end
        results.append(g.run_mixed(threads, rounds, args.seed + k))
    t1 = time.perf_counter()
    for k, r in enumerate(results):
        log(f"[bench] step {k}: {r}")
    lat = g.latency_ns()
    errors = g.errors()

    wall = t1 - t0
    ops = sum(r["mallocs"] + r["frees"] for r in results)
    mallocs = sum(r["mallocs"] for r in results)
    frees = sum(r["frees"] for r in results)
    misses = sum(r["misses"] for r in results)
    tag_errors = sum(r["tag_errors"] for r in results)
    alloc_errors = sum(r["alloc_errors"] for r in results)
    malloc_ms = sum(r["malloc_kernel_ms"] for r in results)
    free_ms = sum(r["free_kernel_ms"] for r in results)
    launches = sum(r["launches"] for r in results)
    kernel_ms = malloc_ms + free_ms

    # dominant kernel roofline (HBM): algorithmic bytes per launch / launch time
    if malloc_ms >= free_ms:
        dom, n_launch, dom_ms, dom_bytes = "mixed_malloc_kernel", rounds * args.steps, malloc_ms, mallocs * BYTES_MALLOC_SIDE
    else:
        dom, n_launch, dom_ms, dom_bytes = "mixed_free_kernel", (rounds + 1) * args.steps, free_ms, frees * BYTES_FREE_SIDE
    achieved = (dom_bytes / n_launch) / ((dom_ms / n_launch) * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    default = (args.pool_gib, args.threads_log2, args.rounds) == DEFAULT_WORKLOAD
    sha = sources_sha()
    current = default and sha == MEASURED["sources_sha"]
    traffic = MEASURED["traffic"][dom] if current else None

    import numpy as np

    p50 = float(np.percentile(lat, 50)) if len(lat) else None
    p99 = float(np.percentile(lat, 99)) if len(lat) else None

    out = {
        "metric": METRIC,
        "value": ops / wall,
        "unit": "mallocs+frees/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (sizes, frees and tags drawn on device from a seed)",
        "config": {
            "workload": f"mixed slice sizes: {args.pool_gib} GiB pool, Gallatin<16MiB,16,4096>, "
            f"{threads} threads/round x {rounds} rounds, sizes 16 B-4 KiB log-uniform, "
            "seeded 50% frees after each round, drain",
            "pool_bytes": pool,
            "threads_per_round": threads,
            "rounds": rounds,
            "parallelism": "single GPU",
        },
        "p50_alloc_latency_ns": p50,
        "p99_alloc_latency_ns": p99,
        "latency_samples": int(len(lat)),
        "mallocs": mallocs,
        "frees": frees,
        "misses": misses,
        "invariant_violations": tag_errors + alloc_errors + errors[0],
        "kernel_ms_per_step": kernel_ms / args.steps,
        "kernel_rate": ops / (kernel_ms * 1e-3) if kernel_ms > 0 else None,
        "launches": launches,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dom_bytes / n_launch,
            "kernel_avg_us": dom_ms / n_launch * 1e3,
            "counters": {"sources_sha": sha, "measured_on_sources": MEASURED["sources_sha"] or None,
                         "commit": MEASURED["commit"] or None, "current": current},
        },
    }
    g.free_on_device()

    # the device-scope atomic rates that bound the slice path
    out["atomic_rates_per_s"] = {
        "distinct_lanes": ga.atomic_rate(ga.ATOMIC_DISTINCT, 1 << 22, 64),
        "wave_hot_4_words": ga.atomic_rate(ga.ATOMIC_WAVE_HOT, 1 << 22, 16, 4),
        "wave_hot_128_words": ga.atomic_rate(ga.ATOMIC_WAVE_HOT, 1 << 22, 16, 128),
        "wave_hot_65536_words": ga.atomic_rate(ga.ATOMIC_WAVE_HOT, 1 << 22, 16, 65536),
    }
    log(f"[bench] atomic rates {out['atomic_rates_per_s']}")
    if current and MEASURED["atomic_requests"][out["roofline"]["kernel"]]:
        # the same kernel against the atomic request rate this GPU sustains on
        # distinct words (the probe's lane rate / 16 lanes per request) and,
        # for one hot line, the single-line rate
        dom = out["roofline"]["kernel"]
        reqs = MEASURED["atomic_requests"][dom]
        achieved_req = reqs / (out["roofline"]["kernel_avg_us"] * 1e-6)
        peak_req = out["atomic_rates_per_s"]["distinct_lanes"] / 16
        out["atomic_roofline"] = {
            "bound": "agent-scope atomics",
            "kernel": dom,
            "achieved": achieved_req,
            "peak": peak_req,
            "unit": "requests/s",
            "frac": achieved_req / peak_req,
            "requests_per_launch": reqs,
            "one_line_peak": out["atomic_rates_per_s"]["wave_hot_4_words"] / 4,
        }

    if not args.headline_only:
        out["plain"] = plain(args.seed, 5, not args.no_cpu_baseline)
        out["forms"] = forms(args.seed)
        out["serving"] = serving(args.seed, args.serving_requests, args.serving_rate)
        # one malloc per page: the same requests, the loop bounded at 5 s past the last arrival
        out["serving_per_page"] = serving(args.seed, args.serving_requests, args.serving_rate, run=1, grace_s=5.0)
        out["blocks"] = blocks(args.seed)
        out["slab"] = slab(args.seed)

    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(pool, args.seed, 1 << args.cpu_threads_log2, rounds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
