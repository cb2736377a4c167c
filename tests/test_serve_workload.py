"""The serving loop's request generator (gal_serve_generate, host code in the
C-ABI library) and the expected-checksum restatement the GPU tests use.

No GPU: gal_serve_generate runs on the host. Statistics are pinned by
closed forms (exponential gaps, lognormal median, the cap's tail mass,
uniform page classes); the exact draws by an independent restatement of the
documented rule in Python floats.
"""
import math

import numpy as np
import pytest

import gallatin_amd as ga
from tests import serving_ref


def unit(seed, i, salt):
    x = serving_ref.splitmix64_int(seed ^ serving_ref.splitmix64_int((i << 4) | salt))
    return ((x >> 11) + 1.0) * 2.0**-53


def test_deterministic_and_ids():
    a = ga.generate_requests(5000, 1000.0, seed=7)
    b = ga.generate_requests(5000, 1000.0, seed=7)
    c = ga.generate_requests(5000, 1000.0, seed=8)
    assert np.array_equal(a, b)
    assert not np.array_equal(a["pages"], c["pages"])
    assert np.array_equal(a["id"], np.arange(5000, dtype=np.uint64))
    assert (a["page_bytes"] == 16384).all()


def test_exact_draws_match_the_documented_rule():
    seed, rate, median, sigma, cap = 99, 50000.0, 1024.0, 1.0, 32768
    r = ga.generate_requests(200, rate, median, sigma, cap, 16384, seed=seed)
    t = 0.0
    for i in range(200):
        t += -math.log(unit(seed, i, 0)) / rate
        z = math.sqrt(-2.0 * math.log(unit(seed, i, 1))) * math.cos(2.0 * math.pi * unit(seed, i, 2))
        length = min(cap, max(1, math.ceil(median * math.exp(sigma * z))))
        assert int(r["pages"][i]) == length, i
        assert abs(int(r["arrival_ns"][i]) - int(t * 1e9)) <= 1, i


def test_poisson_arrivals():
    n, rate = 200000, 50000.0
    r = ga.generate_requests(n, rate, seed=3)
    t = r["arrival_ns"].astype(np.float64) * 1e-9
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert (gaps >= 0).all()
    # exponential gaps: mean 1/rate, standard deviation 1/rate
    assert abs(gaps.mean() * rate - 1.0) < 0.01
    assert abs(gaps.std() * rate - 1.0) < 0.02
    # memorylessness: P(gap > 1/rate) = e^-1
    assert abs((gaps > 1.0 / rate).mean() - math.exp(-1.0)) < 0.005


def test_all_at_once_when_rate_is_zero():
    r = ga.generate_requests(1000, 0.0, seed=4)
    assert (r["arrival_ns"] == 0).all()


def test_lognormal_lengths_and_cap():
    n = 400000
    r = ga.generate_requests(n, 0.0, median_pages=1024, sigma=1.0, max_pages=32768, seed=5)
    pages = r["pages"].astype(np.float64)
    assert pages.min() >= 1 and pages.max() <= 32768
    # median of ceil(1024 e^z) is 1024 (+1 at most from the ceiling)
    assert abs(np.median(pages) - 1024.5) < 20
    # log-lengths: mean ln(1024), sd 1 (ceil and cap shift them slightly)
    logs = np.log(pages[pages < 32768])
    assert abs(logs.mean() - math.log(1024)) < 0.01
    assert abs(logs.std() - 1.0) < 0.01
    # tail mass at the cap: P(z >= ln(32)) = 1 - Phi(3.4657)
    tail = 0.5 * math.erfc(math.log(32.0) / math.sqrt(2.0))
    assert abs((pages == 32768).mean() - tail) < 1e-4


def test_sigma_zero_is_the_median():
    r = ga.generate_requests(100, 0.0, median_pages=37.2, sigma=0.0, seed=6)
    assert (r["pages"] == 38).all()


def test_mixed_page_sizes():
    n = 90000
    r = ga.generate_requests(n, 0.0, page_bytes=0, size_classes=9, seed=7)
    sizes, counts = np.unique(r["page_bytes"], return_counts=True)
    assert list(sizes) == [16 << k for k in range(9)]
    assert (np.abs(counts - n / 9) < 5 * math.sqrt(n / 9)).all()


@pytest.mark.parametrize("kwargs", [dict(page_bytes=8), dict(page_bytes=0, size_classes=0), dict(max_pages=0)])
def test_rejects_bad_workloads(kwargs):
    args = dict(n=10, rate_per_s=1.0)
    args.update(kwargs)
    with pytest.raises(ga.GallatinError):
        ga.generate_requests(**args)


def test_checksum_vectorised_matches_scalar():
    ids = np.array([0, 1, 5, 123456789, 2**63 + 7], dtype=np.uint64)
    pages = np.array([1, 2, 7, 13, 3])
    page_bytes = np.array([16, 32, 4096, 64, 16384])
    got = serving_ref.expected_checksums(ids, pages, page_bytes)
    for i in range(len(ids)):
        assert int(got[i]) == serving_ref.expected_checksum_scalar(int(ids[i]), int(pages[i]), int(page_bytes[i]))


def test_checksum_depends_on_every_read():
    # a plausible slip (a dropped page, another id) changes it; the page size
    # does not enter (only the 16-byte header is written)
    base = serving_ref.expected_checksum_scalar(11, 6, 64)
    assert base != serving_ref.expected_checksum_scalar(11, 5, 64)
    assert base != serving_ref.expected_checksum_scalar(12, 6, 64)
    assert base == serving_ref.expected_checksum_scalar(11, 6, 16) == serving_ref.expected_checksum_scalar(11, 6, 16384)
    # skipping any one of the 2 * pages - 1 reads, or reading one page twice, changes it
    m = (1 << 64) - 1
    word = lambda s: serving_ref.splitmix64_int(serving_ref.splitmix64_int(11) ^ ((s << 3) | 1))  # noqa: E731
    reads = list(range(0, 5)) + list(range(5, -1, -1))  # growth reads pages 0..4, drain 5..0
    assert len(reads) == 2 * 6 - 1

    def checksum(seq):
        h = 11
        for s in seq:
            h = serving_ref.splitmix64_int(h ^ word(s)) & m
        return serving_ref.splitmix64_int(h ^ 6)

    assert checksum(reads) == base
    for j in range(len(reads)):
        assert checksum(reads[:j] + reads[j + 1:]) != base
        assert checksum(reads[:j] + [reads[j]] + reads[j:]) != base
    assert serving_ref.splitmix64_int(0) == 0xE220A8397B1DCDAF  # splitmix64 known answer (Vigna)


def test_checksum_c_matches_numpy():
    rng = np.random.default_rng(3)
    n = 300
    ids = rng.integers(0, 2**63, n, dtype=np.uint64)
    pages = rng.integers(1, 200, n)
    page_bytes = rng.choice([16, 24, 32, 48, 64, 4096, 16384], n)
    expect = serving_ref.expected_checksums(ids, pages, page_bytes)
    got = serving_ref.checksums_c(ids, pages, page_bytes)
    assert np.array_equal(got, expect)
