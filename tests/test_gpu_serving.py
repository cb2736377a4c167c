"""The serving loop on the GPU, through the C-ABI (gal_serve_*).

Every request's checksum must equal the value computed on the host from
(id, pages, page_bytes), and the same request served alone must give the
same checksum (tolerance: none, integer). Admission must hold memory within
the free-segment budget, arrivals must be respected, and after a run the
allocator must be quiescent with its free-segment gauge equal to the
segment tree's population.
"""
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from gallatin_amd import snapshot as snap
from tests import invariants, serving_ref

pytestmark = pytest.mark.gpu

MiB = 1 << 20
SEG = 16 * MiB


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check_after(g):
    assert g.errors() == (0, 0), g.error_sites()
    s = invariants.check_quiescent(g.snapshot(), SEG, 16)
    assert g.free_bytes() == snap.popcount64(s["segment_tree"][-1]) * SEG, "free-segment gauge off"


def check_run(r, reqs, cs, expect_all=True):
    total = int(reqs["pages"].astype(np.int64).sum())
    if expect_all:
        assert r["requests_done"] == len(reqs) and r["requests_aborted"] == 0, r
        assert r["page_mallocs"] == total and r["page_frees"] == total, r
    assert r["alloc_errors"] == 0, r
    exp = serving_ref.expected_checksums(reqs["id"], reqs["pages"], reqs["page_bytes"])
    done = cs != 0
    bad = np.nonzero(done & (cs != exp))[0]
    assert len(bad) == 0, f"{len(bad)} checksums differ, first requests {bad[:8].tolist()}"


def test_free_gauge_tracks_segment_tree():
    g = ga.Gallatin.generate_on_device(1 << 30, 3)
    try:
        s = snap.parse(g.snapshot())
        assert g.free_bytes() == snap.popcount64(s["segment_tree"][-1]) * SEG
        r = g.run_plain(1 << 18, 4096, 5)
        assert r["misses"] == 0
        check_after(g)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("run", [1, 16])
def test_serving_mixed_sizes_alone_equals_mix(run):
    pool = 8 << 30
    reqs = ga.generate_requests(4000, 0.0, median_pages=48, sigma=1.0, max_pages=2048, page_bytes=0,
                                size_classes=9, seed=31)
    say(f"[serve-mix] run={run} n={len(reqs)} pages={int(reqs['pages'].sum())}")
    g = ga.Gallatin.generate_on_device(pool, 12)
    try:
        with ga.Serving(g, len(reqs)) as s:
            s.load(reqs)
            r = s.run(slots=8192, run=run, deadline_s=20.0, window_s=0.002)
            say(f"  mix {r}")
            cs, admit, done = s.results()
            check_run(r, reqs, cs)
            # pages written plus pages freed inside the first 2 ms: some, and at most all of them
            assert 0 < r["window_pages"] <= 2 * int(reqs["pages"].astype(np.int64).sum())
            assert r["first_dry_ms"] > 0
            assert r["peak_active"] > 256, "sequences did not overlap"
            assert (done >= admit).all()
            # the same requests alone
            picks = [int(np.argmax(reqs["pages"])), int(np.argmin(reqs["page_bytes"])), 17, 2500]
            for j in picks:
                one = reqs[j : j + 1].copy()
                s.load(one)
                r1 = s.run(slots=256, run=1 + (j % 7) * run // 4, deadline_s=10.0)
                cs1, _, _ = s.results()
                say(f"  alone {j}: pages {int(one['pages'][0])} x {int(one['page_bytes'][0])} B -> {r1['requests_done']}")
                assert r1["requests_done"] == 1 and int(cs1[0]) == int(cs[j]), (j, int(cs1[0]), int(cs[j]))
        check_after(g)
    finally:
        g.free_on_device()


@pytest.mark.parametrize("run", [1, 32])
def test_serving_admission_gates_on_free_bytes(run):
    pool = 2 << 30
    reqs = ga.generate_requests(300, 0.0, median_pages=512, sigma=1.0, max_pages=4096, page_bytes=16384, seed=41)
    demand = int(reqs["pages"].astype(np.int64).sum()) * 16384
    g = ga.Gallatin.generate_on_device(pool, 13)
    try:
        free0 = g.free_bytes()
        say(f"[serve-admit] demand {demand / 2**30:.2f} GiB, free {free0 / 2**30:.2f} GiB")
        assert demand > free0
        with ga.Serving(g, len(reqs)) as s:
            s.load(reqs)
            r = s.run(slots=4096, run=run, deadline_s=20.0)
            say(f"  run={run} {r}")
            cs, admit, done = s.results()
            check_run(r, reqs, cs)
            assert r["admission_refusals"] > 0, "admission never had to wait"
            assert r["peak_committed"] + r["headroom"] <= free0
            assert r["page_retries"] == 0, "an admitted sequence found no memory"
            assert r["peak_active"] < len(reqs)
        check_after(g)
    finally:
        g.free_on_device()


def test_serving_poisson_arrivals_are_respected_and_overlap():
    pool = 8 << 30
    reqs = ga.generate_requests(3000, 20000.0, median_pages=64, sigma=1.0, max_pages=4096, page_bytes=0,
                                size_classes=9, seed=51)
    say(f"[serve-poisson] n={len(reqs)} span {int(reqs['arrival_ns'][-1]) / 1e6:.1f} ms")
    g = ga.Gallatin.generate_on_device(pool, 14)
    try:
        with ga.Serving(g, len(reqs)) as s:
            s.load(reqs)
            r = s.run(slots=4096, deadline_s=20.0)
            say(f"  {r}")
            cs, admit, done = s.results()
            check_run(r, reqs, cs)
            arrival = reqs["arrival_ns"].astype(np.int64)
            admit, done = admit.astype(np.int64), done.astype(np.int64)
            assert (admit >= arrival - 20).all(), "a request was admitted before it arrived"
            assert (done >= admit).all()
            # continuous batching: requests admitted while others are in flight
            order = np.argsort(admit)
            in_flight = np.searchsorted(np.sort(done), admit[order], side="right")
            concurrent = np.arange(len(order)) - in_flight  # admitted earlier and not yet done
            assert concurrent.max() > 8, "no overlap between sequences"
            assert r["makespan_ms"] * 1e6 >= arrival[-1]
    finally:
        g.free_on_device()


@pytest.mark.parametrize("run", [1, 8])
def test_serving_deadline_aborts_cleanly(run):
    pool = 4 << 30
    reqs = ga.generate_requests(2000, 0.0, median_pages=4096, sigma=0.5, max_pages=32768, page_bytes=16384, seed=61)
    g = ga.Gallatin.generate_on_device(pool, 15)
    try:
        with ga.Serving(g, len(reqs)) as s:
            s.load(reqs)
            r = s.run(slots=1024, run=run, deadline_s=0.002)
            say(f"[serve-deadline] run={run} {r}")
            cs, admit, done = s.results()
            assert r["requests_done"] + r["requests_aborted"] == len(reqs)
            assert r["requests_aborted"] > 0
            assert r["page_frees"] == r["page_mallocs"]
            cut = done == np.uint64(2**64 - 1)
            assert int(cut.sum()) == r["requests_aborted"] and (cs[cut] == 0).all()
            check_run(r, reqs, cs, expect_all=False)
        check_after(g)
    finally:
        g.free_on_device()


def test_serving_rejects_a_request_that_never_fits():
    g = ga.Gallatin.generate_on_device(1 << 30, 16)
    try:
        reqs = ga.generate_requests(4, 0.0, median_pages=1 << 16, sigma=0.0, max_pages=1 << 16, page_bytes=16384)
        with ga.Serving(g, 4) as s:
            s.load(reqs)
            with pytest.raises(ga.GallatinError):
                s.run(slots=256)
        check_after(g)
    finally:
        g.free_on_device()


def test_serving_invalid_load_keeps_the_previous_one():
    """A load with an invalid request (no pages) is refused before anything
    of the previous load changes: that load still runs, with every checksum
    right."""
    # 8 GiB: the default admission headroom of nine page classes (their
    # slots' blocks and two segments each) needs about 1 GiB
    g = ga.Gallatin.generate_on_device(8 << 30, 17)
    try:
        reqs = ga.generate_requests(200, 0.0, median_pages=8, sigma=1.0, max_pages=64, page_bytes=0, seed=17)
        bad = reqs.copy()
        bad["pages"][3] = 0
        with ga.Serving(g, len(reqs)) as s:
            s.load(reqs)
            with pytest.raises(ga.GallatinError):
                s.load(bad)
            r = s.run(slots=256, run=8)
            cs, _, _ = s.results()
            check_run(r, reqs, cs)
        check_after(g)
    finally:
        g.free_on_device()
