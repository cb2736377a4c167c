"""Expected results of the serving loop, computed on the host with numpy.

A sequence of `pages` pages (of any size: only the 16-byte header is
written), request id `id`:
 * page s holds the content word page_word(id, s) (word 0 links to the
   previous page);
 * growing: after page s >= 1 is allocated, the word of page s-1 is read
   back and folded into the checksum h (h starts at id);
 * draining: pages L-1 down to 0 are read back and folded, then freed;
 * the result is fold(h, pages).
fold(h, x) = splitmix64(h ^ x); page_word(id, s) = splitmix64(splitmix64(id) ^ (s << 3 | 1)).
Vectorised over requests: the loop runs over page index.
"""
import numpy as np

_C1 = np.uint64(0x9E3779B97F4A7C15)
_C2 = np.uint64(0xBF58476D1CE4E5B9)
_C3 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + _C1
        z = (z ^ (z >> np.uint64(30))) * _C2
        z = (z ^ (z >> np.uint64(27))) * _C3
    return z ^ (z >> np.uint64(31))


def splitmix64_int(x):
    """Scalar reference on Python ints (independent of numpy's wrapping)."""
    m = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def expected_checksums(ids, pages, page_bytes):
    ids = np.asarray(ids, dtype=np.uint64)
    pages = np.asarray(pages, dtype=np.int64)
    assert np.all(np.asarray(page_bytes) >= 16)
    base = splitmix64(ids)
    h = ids.copy()
    top = int(pages.max()) if len(pages) else 0

    def fold_page(h, s, active):
        s64 = np.asarray(s, dtype=np.uint64)
        w = splitmix64(base ^ ((s64 << np.uint64(3)) | np.uint64(1)))
        return np.where(active, splitmix64(h ^ w), h)

    # growth: page s-1 read back after page s is allocated, s = 1 .. pages-1
    for s in range(1, top):
        h = fold_page(h, s - 1, pages > s)
    # drain: page s read back for s = pages-1 .. 0
    for s in range(top - 1, -1, -1):
        h = fold_page(h, s, pages > s)
    return splitmix64(h ^ pages.astype(np.uint64))


def expected_checksum_scalar(rid, pages, page_bytes):
    """The same on Python ints, one request (small cases only)."""
    assert page_bytes >= 16
    word = lambda s: splitmix64_int(splitmix64_int(rid) ^ ((s << 3) | 1))  # noqa: E731
    h = rid
    for s in range(1, pages):
        h = splitmix64_int(h ^ word(s - 1))
    for s in range(pages - 1, -1, -1):
        h = splitmix64_int(h ^ word(s))
    return splitmix64_int(h ^ pages)


def checksums_c(ids, pages, page_bytes):
    """The same computed by tests/native/serving_check.c (large samples)."""
    import ctypes
    import os

    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "build",
                                   "libserving_check.so"))
    ids = np.ascontiguousarray(ids, dtype=np.uint64)
    pages = np.ascontiguousarray(pages, dtype=np.uint32)
    page_bytes = np.ascontiguousarray(page_bytes, dtype=np.uint32)
    out = np.zeros(len(ids), dtype=np.uint64)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    lib.serving_checksums(p(ids), p(pages), p(page_bytes), ctypes.c_uint64(len(ids)), p(out))
    return out
