"""Autotuner logic without a GPU: legality of instantiations, the candidate
set, the choice made from measured trials, and the result cache. The trial
measurements are stand-ins passed as `measure`; the GPU trial itself is
covered by tests/test_gpu_slab.py."""
import pytest

from gallatin_amd import autotune

MiB = 1 << 20


@pytest.mark.parametrize(
    "triple, ok",
    [
        ((4 * MiB, 128, 1024), True),
        ((4 * MiB, 128, 4096), False),  # a segment must hold one block of the top class
        ((16 * MiB, 16, 4096), True),
        ((16 * MiB, 24, 4096), False),  # not a power of two
        ((1 << 40, 1, 1 << 31), False),  # more than 30 classes
    ],
)
def test_legal(triple, ok):
    assert autotune.legal(*triple) == ok


def test_candidates_are_the_compiled_min128_triples():
    c = autotune.candidates(128)
    assert c == sorted(c) and len(c) >= 3
    assert all(mn == 128 and autotune.legal(seg, mn, mx) for seg, mn, mx in c)
    assert {seg for seg, _, _ in c} >= {4 * MiB, 16 * MiB, 64 * MiB}
    assert autotune.candidates(128, segment_sizes=(16 * MiB,)) == [(16 * MiB, 128, 4096)]


def fake_measure(rates, calls=None):
    """Trials whose batch rate per (width, run) and slab rate come from `rates`;
    only the (width, run) settings asked for are measured."""

    def measure(triple, pool, slab_keys, slab_buckets, batch_requests, classes, replays, widths, runs):
        if calls is not None:
            calls.append((triple, tuple(widths), tuple(runs)))
        slab_rate, batch = rates[triple]
        asked = {(w, r) for w in widths for r in runs}
        return {
            "slab": {"ops": slab_keys, "ms": slab_keys / slab_rate * 1e3},
            "batch": {wr: {"ops": batch_requests * 2, "ms": batch_requests * 2 / r * 1e3}
                      for wr, r in batch.items() if wr in asked},
        }

    return measure


def test_picks_best_width_run_per_triple_and_best_triple(tmp_path):
    c = autotune.candidates(128)
    # the probe setting (last width, last run) = (64, 16)
    rates = {t: (1e9, {(64, 16): 1e9, (32, 16): 2e9}) for t in c}
    best_t = c[1]
    rates[best_t] = (1e9, {(64, 16): 1.5e9, (16, 4): 4e9, (8, 1): 3e9})
    rates[c[0]] = (1e9, {(64, 16): 0.5e9, (8, 1): 9e9})  # probes worst: never gets the grid
    calls = []
    out = autotune.tune(1 << 30, slab_keys=1000, batch_requests=1000, measure=fake_measure(rates, calls))
    assert out["triple"] == best_t and (out["width"], out["run"]) == (16, 4)
    # combined rate: 1000 slab ops at 1e9/s and 2000 batch ops at 4e9/s
    assert out["rate"] == pytest.approx(3000 / (1000 / 1e9 + 2000 / 4e9))
    probes = [x for x in calls if x[1] == (64,) and x[2] == (16,)]
    grids = [x for x in calls if x not in probes]
    assert len(probes) == len(c) and len(grids) == 2 and c[0] not in {g[0] for g in grids}

    cache = str(tmp_path / "tune.json")
    a = autotune.tune(1 << 30, slab_keys=1000, batch_requests=1000, measure=fake_measure(rates), cache=cache)

    def fail(*args):
        raise AssertionError("cache miss")

    b = autotune.tune(1 << 30, slab_keys=1000, batch_requests=1000, measure=fail, cache=cache)
    assert b["triple"] == a["triple"] and b["width"] == a["width"] and b["run"] == a["run"]


def test_a_triple_whose_trials_all_missed_is_skipped():
    c = autotune.candidates(128)
    rates = {t: (1e9, {(64, 16): 1e9}) for t in c}
    rates[c[0]] = (1e12, {})  # every batch trial missed
    out = autotune.tune(1 << 30, slab_keys=10, batch_requests=10, measure=fake_measure(rates))
    assert out["triple"] != c[0]
    assert any(t.get("note") for t in out["trials"])


def test_rejects_unknown_min_size():
    with pytest.raises(ValueError):
        autotune.tune(1 << 30, min_size=96)
