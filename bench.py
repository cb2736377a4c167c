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
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "device mallocs+frees/sec on 1 MI355X (mixed sizes); p99 alloc latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per malloc+free pair on the slice path: metadata 26 B
# (8 B pinned-block pointer read, 4+4 B malloc-counter RMW, 4+4 B free-counter
# RMW, 2 B chunk-id read) + payload 16 B (8 B tag write, 8 B tag read).
# Malloc kernel share: 8 + 8 + 8 = 24 B; free kernel share: 8 + 2 + 8 = 18 B.
BYTES_MALLOC_SIDE = 24
BYTES_FREE_SIDE = 18
# HBM bytes per launch of the workload kernels at the default workload, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE runs (KiB per dispatch, averaged over a
# step) scaled by what the counters read for a 1 GiB device copy (fetch x2.0,
# write x1.0): profiles/mixed_slices_8GiB_hbm_traffic.md.
MEASURED_TRAFFIC_BYTES = {"mixed_malloc_kernel": 212.8e6, "mixed_free_kernel": 395.0e6}
DEFAULT_WORKLOAD = (8, 22, 10)  # pool GiB, log2 threads per round, rounds


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
    ap.add_argument("--cpu-threads-log2", type=int, default=22, help="oracle sample size (threads per round)")
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

    def step(seed, tag):
        """One step: reset, every round (malloc kernel + free kernel), drain."""
        g.mixed_reset(threads, rounds, seed)
        total = None
        before = g.stats()
        for r in range(rounds):
            res = g.mixed_round(r)
            now = g.stats()
            delta = {k: now[k] - before[k] for k in now if now[k] != before[k]}
            before = now
            log(f"[bench] {tag} round {r}: mallocs {res['mallocs']} frees {res['frees']} misses {res['misses']} "
                f"malloc_ms {res['malloc_kernel_ms']:.2f} free_ms {res['free_kernel_ms']:.2f} rare paths {delta}")
            total = res if total is None else {k: total[k] + res[k] for k in total}
        res = g.mixed_drain()
        total = {k: total[k] + res[k] for k in total}
        return total

    for w in range(args.warmup):
        r = step(args.seed + 1000 + w, f"warmup {w}")
        log(f"[bench] warmup {w}: {r}")

    results = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        results.append(step(args.seed + k, f"step {k}"))
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
    traffic = MEASURED_TRAFFIC_BYTES[dom] if default else None

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

    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        out["cpu_baseline"] = cpu_baseline(pool, args.seed, 1 << args.cpu_threads_log2, rounds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
