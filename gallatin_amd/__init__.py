"""MI355X-native Gallatin: a device-side malloc/free for gfx950.

Host mirror of the reference's host interface (include/gallatin/allocators/
gallatin.cuh:324-421, 643-668, 1855-1897 and global_allocator.cuh:48-97),
backed by the HIP library gallatin_amd/lib/libgallatin_amd.so. The device
interface (Gallatin<SEG,MIN,MAX>::malloc/free, global_malloc/global_free) is
the C++ header set in gallatin_amd/include/gallatin/. The workloads that
drive an allocator (run_plain, the mixed rounds, churn, batches, the serving
loop, the slab table, the global-API kernels) come from the separate library
gallatin_amd/lib/libgallatin_workloads.so, each through a workload object
built on the allocator's handle.
"""
import ctypes
import weakref

import numpy as np

from . import _capi
from ._capi import (  # noqa: F401
    GAL_NULL_OFFSET,
    GAL_OP_FREE,
    GAL_OP_MALLOC,
    GAL_VERIFY_ATOMIC,
    GAL_VERIFY_STORE,
    GallatinError,
    GalGroupOp,
    GalOp,
)

DEFAULT_SEGMENT = 16 << 20
DEFAULT_MIN = 16
DEFAULT_MAX = 4096

_lib = _capi.lib
_wl = _capi.wl


def _new_workload(h):
    """A workload object (libgallatin_workloads.so) on allocator handle h."""
    w = ctypes.c_void_p()
    _capi.check("gal_workload_create", _wl.gal_workload_create(h, ctypes.byref(w)))
    return w


def supported_triples():
    buf = (ctypes.c_uint64 * 96)()
    n = _lib.gal_supported(buf, 32)
    return [tuple(buf[3 * i : 3 * i + 3]) for i in range(n)]


def ops_array(ops):
    """ops: iterable of (kind, arg, tid, slot) -> ctypes array of gal_op."""
    ops = list(ops)
    arr = (GalOp * max(1, len(ops)))()
    for i, (kind, arg, tid, slot) in enumerate(ops):
        arr[i].kind = kind
        arr[i].arg = arg
        arr[i].tid = tid
        arr[i].slot = slot
    return arr, len(ops)


def group_ops_array(ops):
    """ops: iterable of (kind, slot, lanes, args, tids) with args/tids indexed
    by lane (64 entries, or a dict lane -> value) -> ctypes array of gal_group_op."""
    ops = list(ops)
    arr = (GalGroupOp * max(1, len(ops)))()
    for i, (kind, slot, lanes, args, tids) in enumerate(ops):
        arr[i].kind = kind
        arr[i].slot = slot
        arr[i].lanes = lanes
        for lane in range(64):
            if (lanes >> lane) & 1:
                arr[i].arg[lane] = int(args[lane])
                arr[i].tid[lane] = int(tids[lane])
    return arr, len(ops)


class Gallatin:
    """Host handle of one device allocator Gallatin<segment_size, min_size, max_size>."""

    def __init__(self, handle, info, owned=True):
        self._h = handle
        self.info = info
        self._owned = owned
        self._children = []  # weak references to batches, serving loops, slab tables made on it
        self._w = None  # this handle's workload object, made on first use

    def _workload(self):
        if self._w is None:
            self._w = _new_workload(self._h)
        return self._w

    def _drop_workload(self):
        """Destroy the workload object (and its hold on the handle) unless a
        batch, serving loop or slab table still uses it."""
        if self._w is not None and _wl.gal_workload_destroy(self._w) == _capi.GAL_OK:
            self._w = None

    def _adopt(self, child):
        self._children = [c for c in self._children if c() is not None]
        self._children.append(weakref.ref(child))

    # reference gallatin.cuh:324 generate_on_device(max_bytes, seed, print_info, running_calloc)
    @classmethod
    def generate_on_device(
        cls,
        max_bytes,
        seed,
        print_info=False,
        segment_size=DEFAULT_SEGMENT,
        min_size=DEFAULT_MIN,
        max_size=DEFAULT_MAX,
    ):
        h = ctypes.c_void_p()
        _capi.check("gal_create", _lib.gal_create(segment_size, min_size, max_size, max_bytes, seed, ctypes.byref(h)))
        info = _capi.GalInfo()
        _capi.check("gal_get_info", _lib.gal_get_info(h, ctypes.byref(info)))
        obj = cls(h, info)
        if print_info:
            obj.print_info()
        return obj

    # reference gallatin.cuh:427 generate_on_device_host: the pool is mapped host memory
    @classmethod
    def generate_on_device_host(
        cls,
        max_bytes,
        seed,
        print_info=False,
        segment_size=DEFAULT_SEGMENT,
        min_size=DEFAULT_MIN,
        max_size=DEFAULT_MAX,
    ):
        h = ctypes.c_void_p()
        _capi.check("gal_create_host",
                    _lib.gal_create_host(segment_size, min_size, max_size, max_bytes, seed, ctypes.byref(h)))
        info = _capi.GalInfo()
        _capi.check("gal_get_info", _lib.gal_get_info(h, ctypes.byref(info)))
        obj = cls(h, info)
        if print_info:
            obj.print_info()
        return obj

    # reference gallatin.cuh:643 free_on_device
    def free_on_device(self):
        # the library refuses to destroy a handle that batches, serving loops
        # or slab tables still use (GAL_ERR_STATE): they are released first
        for ref in self._children:
            child = ref()
            if child is not None:
                child._release()
        self._children = []
        if self._w is not None:
            _capi.check("gal_workload_destroy", _wl.gal_workload_destroy(self._w))
            self._w = None
        if self._h is not None and self._owned:
            _capi.check("gal_destroy", _lib.gal_destroy(self._h))
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free_on_device()

    def __del__(self):
        try:
            self.free_on_device()
        except Exception:
            pass

    # reference gallatin.cuh:1855 print_info
    def print_info(self):
        _capi.check("gal_print_stats", _lib.gal_print_stats(self._h))

    def fill(self):
        """print_info's figures as numbers (gal_fill): per class lists of length num_trees."""
        r = _capi.GalFillReport()
        _capi.check("gal_fill", _lib.gal_fill(self._h, ctypes.byref(r)))
        t = r.num_trees
        return {
            "free_segments": r.free_segments,
            "max_segments": r.max_segments,
            "listed_segments": list(r.listed_segments[:t]),
            "live_segments": list(r.live_segments[:t]),
            "blocks_available": list(r.blocks_available[:t]),
            "blocks_in_use": list(r.blocks_in_use[:t]),
            "overhead_bytes": r.overhead_bytes,
            "empty_pinned_slots": [int(ctypes.c_int64(x).value) for x in r.empty_pinned_slots[:t]],
        }

    @property
    def handle(self):
        return self._h

    def errors(self):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        _capi.check("gal_errors", _lib.gal_errors(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    ERROR_SITES = (
        "queue_empty",
        "reset_count",
        "chunk_reset",
        "setup",
        "queue_spin",
        "final_spin",
        "fill_class",
        "run_reset",
        "fill_count",
        "init_counter",
        "grant_class",
        "claim_class",
        "free_class",
        "over_return",
    )

    def error_sites(self):
        buf = (ctypes.c_uint32 * 14)()
        _capi.check("gal_error_sites", _lib.gal_error_sites(self._h, buf, 14))
        return {name: buf[i] for i, name in enumerate(self.ERROR_SITES) if buf[i]}

    STAT_NAMES = (
        "replace",
        "refill",
        "request",
        "request_retry",
        "attach",
        "release",
        "recycle",
        "replace_ticks16",
        "refill_ticks16",
        "ahead",
        "ahead_taken",
    )

    def stats(self):
        n = len(self.STAT_NAMES)
        buf = (ctypes.c_uint32 * n)()
        _capi.check("gal_stats", _lib.gal_stats(self._h, buf, n))
        return {name: buf[i] for i, name in enumerate(self.STAT_NAMES)}

    def snapshot(self):
        need = ctypes.c_uint64()
        _lib.gal_snapshot(self._h, None, 0, ctypes.byref(need))
        buf = np.zeros(need.value, dtype=np.uint8)
        _capi.check(
            "gal_snapshot", _lib.gal_snapshot(self._h, buf.ctypes.data_as(ctypes.c_void_p), need.value, None)
        )
        return buf

    def replay(self, ops):
        arr, n = ops_array(ops)
        out = (ctypes.c_uint64 * max(1, n))()
        _capi.check("gal_replay", _lib.gal_replay(self._h, arr, n, out))
        return np.frombuffer(out, dtype=np.uint64, count=n).copy()

    def group_replay(self, ops):
        """Multi-lane deterministic replay (gal_group_replay); returns an
        (n, 64) array of per-lane results."""
        arr, n = group_ops_array(ops)
        out = (ctypes.c_uint64 * max(1, 64 * n))()
        _capi.check("gal_group_replay", _lib.gal_group_replay(self._h, arr, n, out))
        return np.frombuffer(out, dtype=np.uint64, count=64 * n).copy().reshape(n, 64)

    def group_replay_noalloc(self, d_ops, n, d_results):
        """gal_group_replay_noalloc on device buffers the caller owns (device
        addresses as ints); returns the status code (0, or GAL_ERR_*)."""
        return _lib.gal_group_replay_noalloc(self._h, d_ops, n, d_results)

    def run_plain(self, n_threads, size, seed, verify=GAL_VERIFY_STORE):
        r = _capi.GalRunResult()
        _capi.check("gal_run_plain", _wl.gal_run_plain(self._workload(), n_threads, size, seed, verify, ctypes.byref(r)))
        return r.as_dict()

    def mixed_reset(self, threads_per_round, rounds, seed):
        _capi.check("gal_mixed_reset", _wl.gal_mixed_reset(self._workload(), threads_per_round, rounds, seed))

    def mixed_round(self, rnd, verify=GAL_VERIFY_STORE):
        r = _capi.GalRunResult()
        _capi.check("gal_mixed_round", _wl.gal_mixed_round(self._workload(), rnd, verify, ctypes.byref(r)))
        return r.as_dict()

    def mixed_drain(self, verify=GAL_VERIFY_STORE):
        r = _capi.GalRunResult()
        _capi.check("gal_mixed_drain", _wl.gal_mixed_drain(self._workload(), verify, ctypes.byref(r)))
        return r.as_dict()

    def mixed_live(self):
        n = ctypes.c_uint64()
        _capi.check("gal_mixed_live", _wl.gal_mixed_live(self._workload(), None, None, None, 0, ctypes.byref(n)))
        cnt = n.value
        offs = np.zeros(max(1, cnt), dtype=np.uint64)
        sizes = np.zeros(max(1, cnt), dtype=np.uint64)
        slots = np.zeros(max(1, cnt), dtype=np.uint64)
        p = ctypes.POINTER(ctypes.c_uint64)
        _capi.check(
            "gal_mixed_live",
            _wl.gal_mixed_live(
                self._workload(),
                offs.ctypes.data_as(p),
                sizes.ctypes.data_as(p),
                slots.ctypes.data_as(p),
                cnt,
                ctypes.byref(n),
            ),
        )
        return offs[:cnt], sizes[:cnt], slots[:cnt]

    def run_mixed(self, threads_per_round, rounds, seed, verify=GAL_VERIFY_STORE):
        r = _capi.GalRunResult()
        _capi.check(
            "gal_run_mixed", _wl.gal_run_mixed(self._workload(), threads_per_round, rounds, seed, verify, ctypes.byref(r))
        )
        return r.as_dict()

    def latency_keep(self, runs):
        """Keep the latency samples of the next `runs` run_mixed calls (of one
        shape) for latency_ns, instead of the last run's; 0 restores that."""
        _capi.check("gal_latency_keep", _wl.gal_latency_keep(self._workload(), runs))

    def latency_ns(self):
        """Per-wave malloc latencies (ns) of the last mixed run, or of the
        runs kept since latency_keep."""
        n, tick = ctypes.c_uint64(), ctypes.c_double()
        _capi.check("gal_latency_samples", _wl.gal_latency_samples(self._workload(), None, 0, ctypes.byref(n), ctypes.byref(tick)))
        buf = np.zeros(max(1, n.value), dtype=np.uint32)
        _capi.check(
            "gal_latency_samples",
            _wl.gal_latency_samples(
                self._workload(), buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n.value, ctypes.byref(n), None
            ),
        )
        return buf[: n.value].astype(np.float64) * tick.value

    def churn(self, kinds, args, seed=0):
        """One concurrent kernel: op i is malloc(args[i]) (kind GAL_OP_MALLOC) or
        free(pool + args[i]) (GAL_OP_FREE). Returns (results, run result)."""
        kinds = np.asarray(kinds, dtype=np.uint32)
        args = np.asarray(args, dtype=np.uint64)
        n = len(kinds)
        arr = (GalOp * max(1, n))()
        view = np.frombuffer(arr, dtype=np.dtype([("kind", "<u4"), ("slot", "<u4"), ("arg", "<u8"), ("tid", "<u8")]),
                             count=n)
        view["kind"] = kinds
        view["arg"] = args
        out = np.zeros(max(1, n), dtype=np.uint64)
        r = _capi.GalRunResult()
        _capi.check(
            "gal_churn",
            _wl.gal_churn(self._workload(), arr, n, seed, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(r)),
        )
        return out[:n], r.as_dict()

    def largest_segment_run(self):
        """Most contiguous segments one malloc could claim now (gal_largest_run)."""
        out = ctypes.c_uint64()
        _capi.check("gal_largest_run", _lib.gal_largest_run(self._h, ctypes.byref(out)))
        return out.value

    def free_bytes(self):
        """Bytes in free segments (the allocator's gauge, gal_free_bytes)."""
        out = ctypes.c_uint64()
        _capi.check("gal_free_bytes", _lib.gal_free_bytes(self._h, ctypes.byref(out)))
        return out.value

    def read_pool(self, offset, nbytes):
        buf = np.zeros(nbytes, dtype=np.uint8)
        _capi.check(
            "gal_read_pool", _lib.gal_read_pool(self._h, offset, buf.ctypes.data_as(ctypes.c_void_p), nbytes)
        )
        return buf


# ---- global allocator (reference global_allocator.cuh:41-97)
_global_views = weakref.WeakSet()  # Gallatin views handed out by global_handles()


def _drop_view_workloads():
    # a view's idle workload object holds the global handle; the free calls
    # below would refuse it (GAL_ERR_STATE), so it goes first
    for v in list(_global_views):
        v._drop_workload()


def init_global_allocator(num_bytes, seed, print_info=True):
    _capi.check("gal_init_global_allocator", _lib.gal_init_global_allocator(num_bytes, seed, int(bool(print_info))))


def free_global_allocator():
    _drop_view_workloads()
    _capi.check("gal_free_global_allocator", _lib.gal_free_global_allocator())


def print_global_stats():
    _capi.check("gal_print_global_stats", _lib.gal_print_global_stats())


def allocated_size(size, segment_size=DEFAULT_SEGMENT, min_size=DEFAULT_MIN, max_size=DEFAULT_MAX):
    """Bytes malloc grants a request of `size` (gal_allocated_size; host only)."""
    out = ctypes.c_uint64()
    _capi.check("gal_allocated_size",
                _lib.gal_allocated_size(segment_size, min_size, max_size, size, ctypes.byref(out)))
    return out.value


# reference global_allocator.cuh:98-236: a global allocator over host memory, and both
def init_global_allocator_host(num_bytes, seed, print_info=True):
    _capi.check("gal_init_global_allocator_host",
                _lib.gal_init_global_allocator_host(num_bytes, seed, int(bool(print_info))))


def free_global_allocator_host():
    _drop_view_workloads()
    _capi.check("gal_free_global_allocator_host", _lib.gal_free_global_allocator_host())


def print_global_stats_host():
    _capi.check("gal_print_global_stats_host", _lib.gal_print_global_stats_host())


def init_global_allocator_combined(num_bytes, host_bytes, seed, print_info=True):
    _capi.check("gal_init_global_allocator_combined",
                _lib.gal_init_global_allocator_combined(num_bytes, host_bytes, seed, int(bool(print_info))))


def free_global_allocator_combined():
    _drop_view_workloads()
    _capi.check("gal_free_global_allocator_combined", _lib.gal_free_global_allocator_combined())


def print_global_stats_combined():
    _capi.check("gal_print_global_stats_combined", _lib.gal_print_global_stats_combined())


def global_handles():
    """(device, host) global allocators as non-owning Gallatin views (None if absent)."""
    d, h = ctypes.c_void_p(), ctypes.c_void_p()
    _capi.check("gal_global_handles", _lib.gal_global_handles(ctypes.byref(d), ctypes.byref(h)))
    out = []
    for x in (d, h):
        if not x.value:
            out.append(None)
            continue
        info = _capi.GalInfo()
        _capi.check("gal_get_info", _lib.gal_get_info(x, ctypes.byref(info)))
        view = Gallatin(x, info, owned=False)
        _global_views.add(view)
        out.append(view)
    return tuple(out)


def global_fused(n_threads, size):
    """global_malloc_fused / global_free_fused from n threads: (on_device, on_host, misses, tag_errors)."""
    a, b, c, d = (ctypes.c_uint64() for _ in range(4))
    _capi.check("gal_global_fused", _wl.gal_global_fused(n_threads, size, ctypes.byref(a), ctypes.byref(b),
                                                          ctypes.byref(c), ctypes.byref(d)))
    return a.value, b.value, c.value, d.value


def global_poison(size):
    """Guard-check results of the reference's poison sequence for one size (gal_global_poison)."""
    codes = (ctypes.c_int32 * 7)()
    _capi.check("gal_global_poison", _wl.gal_global_poison(size, codes, 7))
    return list(codes)


def global_plain(n_threads, size, seed, verify=GAL_VERIFY_STORE):
    """n_threads threads each global_malloc(size), tag, check, global_free (timed)."""
    r = _capi.GalRunResult()
    _capi.check("gal_global_plain", _wl.gal_global_plain(n_threads, size, seed, verify, ctypes.byref(r)))
    return r.as_dict()


def global_replay(ops):
    """Single-lane replay through global_malloc/global_free; returns (results,
    cu_id): cu_id is the slot hint of the CU the lane ran on (its XCD in the
    low three bits), the slot the same stream replays with."""
    arr, n = ops_array(ops)
    out = (ctypes.c_uint64 * max(1, n))()
    cu = ctypes.c_uint32()
    _capi.check("gal_global_replay", _lib.gal_global_replay(arr, n, out, ctypes.byref(cu)))
    return np.frombuffer(out, dtype=np.uint64, count=n).copy(), cu.value


def global_smoke(n_threads, size):
    m, t = ctypes.c_uint64(), ctypes.c_uint64()
    _capi.check("gal_global_smoke", _wl.gal_global_smoke(n_threads, size, ctypes.byref(m), ctypes.byref(t)))
    return m.value, t.value


# ---- batches of uneven requests (no reference counterpart; same device API)
FORM_PER_THREAD = _capi.GAL_FORM_PER_THREAD
FORM_BATCHED = _capi.GAL_FORM_BATCHED
FORM_GRAPH = _capi.GAL_FORM_GRAPH


class Batch:
    """Requests served together: binned by class on the device, then one
    persistent launch (see gal_batch_* in include/gallatin_workloads.h)."""

    def __init__(self, allocator, capacity):
        self._w = _new_workload(allocator._h)
        h = ctypes.c_void_p()
        rc = _wl.gal_batch_create(self._w, capacity, ctypes.byref(h))
        if rc != _capi.GAL_OK:
            _wl.gal_workload_destroy(self._w)
            _capi.check("gal_batch_create", rc)
        self._b = h
        self.allocator = allocator
        self.capacity = capacity
        self.n = 0
        allocator._adopt(self)

    def close(self):
        if self._b:
            _capi.check("gal_batch_destroy", _wl.gal_batch_destroy(self._b))
            self._b = None
        if self._w:
            _capi.check("gal_workload_destroy", _wl.gal_workload_destroy(self._w))
            self._w = None

    _release = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def generate(self, n, seed, size_classes=9):
        _capi.check("gal_batch_generate", _wl.gal_batch_generate(self._b, n, seed, size_classes))
        self.n = n

    def set_sizes(self, sizes, seed=0):
        arr = np.ascontiguousarray(sizes, dtype=np.uint64)
        p = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        _capi.check("gal_batch_set_sizes", _wl.gal_batch_set_sizes(self._b, p, len(arr), seed))
        self.n = len(arr)

    def run(self, form, width=64, replays=1, verify=GAL_VERIFY_STORE, run=1):
        r = _capi.GalRunResult()
        _capi.check(
            "gal_batch_run", _wl.gal_batch_run(self._b, form, width, run, replays, verify, ctypes.byref(r))
        )
        return r.as_dict()

    def malloc(self, width=64, run=1):
        r = _capi.GalRunResult()
        _capi.check("gal_batch_malloc", _wl.gal_batch_malloc(self._b, width, run, ctypes.byref(r)))
        return r.as_dict()

    def free(self):
        r = _capi.GalRunResult()
        _capi.check("gal_batch_free", _wl.gal_batch_free(self._b, ctypes.byref(r)))
        return r.as_dict()

    def offsets(self):
        out = np.zeros(max(1, self.n), dtype=np.uint64)
        p = out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        _capi.check("gal_batch_offsets", _wl.gal_batch_offsets(self._b, p, self.n))
        return out[: self.n]

    def order(self):
        out = np.zeros(max(1, self.n), dtype=np.uint32)
        p = out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        _capi.check("gal_batch_order", _wl.gal_batch_order(self._b, p, self.n))
        return out[: self.n]


def gather_limit(words):
    """Most contiguous items the segment tree's gather rule could claim from
    these item-layer words (gal_gather_limit; host only)."""
    w = np.ascontiguousarray(words, dtype=np.uint64)
    out = ctypes.c_uint64()
    _capi.check("gal_gather_limit", _lib.gal_gather_limit(w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(w),
                                                          ctypes.byref(out)))
    return out.value


# ---- serving loop (no reference counterpart; every page is a Gallatin malloc/free)
REQUEST_DTYPE = np.dtype([("arrival_ns", "<u8"), ("id", "<u8"), ("pages", "<u4"), ("page_bytes", "<u4")])
SERVE_DEFAULT_HEADROOM = _capi.GAL_SERVE_DEFAULT_HEADROOM


def generate_requests(n, rate_per_s, median_pages=1024.0, sigma=1.0, max_pages=32768, page_bytes=16384,
                      size_classes=9, seed=0):
    """Sequence requests (gal_serve_generate): Poisson arrivals at rate_per_s
    (0: all at time 0), lognormal lengths in pages, fixed page_bytes or, with
    page_bytes=0, 16 << k bytes with k uniform in [0, size_classes)."""
    w = _capi.GalServeWorkload(n, float(rate_per_s), float(median_pages), float(sigma), max_pages, page_bytes,
                               size_classes, 0, seed)
    out = np.zeros(n, dtype=REQUEST_DTYPE)
    _capi.check("gal_serve_generate", _wl.gal_serve_generate(ctypes.byref(w), out.ctypes.data_as(ctypes.c_void_p)))
    return out


class Serving:
    """A device-resident serving loop over one allocator (gal_serve_* in
    include/gallatin_workloads.h): sequences arrive, are admitted by free pool
    bytes, grow one page per step and free their pages when done."""

    def __init__(self, allocator, capacity):
        self._w = _new_workload(allocator._h)
        h = ctypes.c_void_p()
        rc = _wl.gal_serve_create(self._w, capacity, ctypes.byref(h))
        if rc != _capi.GAL_OK:
            _wl.gal_workload_destroy(self._w)
            _capi.check("gal_serve_create", rc)
        self._s = h
        self.allocator = allocator
        self.n = 0
        allocator._adopt(self)

    def close(self):
        if self._s:
            _capi.check("gal_serve_destroy", _wl.gal_serve_destroy(self._s))
            self._s = None
        if self._w:
            _capi.check("gal_workload_destroy", _wl.gal_workload_destroy(self._w))
            self._w = None

    _release = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def load(self, requests):
        arr = np.ascontiguousarray(requests, dtype=REQUEST_DTYPE)
        _capi.check("gal_serve_load", _wl.gal_serve_load(self._s, arr.ctypes.data_as(ctypes.c_void_p), len(arr)))
        self.n = len(arr)

    def run(self, slots=16384, run=1, headroom=SERVE_DEFAULT_HEADROOM, deadline_s=60.0, window_s=0.0):
        r = _capi.GalServeResult()
        _capi.check("gal_serve_run",
                    _wl.gal_serve_run(self._s, slots, run, headroom, deadline_s, window_s, ctypes.byref(r)))
        return r.as_dict()

    def results(self):
        """(checksums, admit_ns, done_ns) per request of the last run; an
        aborted request has checksum 0 and times 2**64-1."""
        n = max(1, self.n)
        out = [np.zeros(n, dtype=np.uint64) for _ in range(3)]
        p = ctypes.POINTER(ctypes.c_uint64)
        _capi.check(
            "gal_serve_results",
            _wl.gal_serve_results(self._s, out[0].ctypes.data_as(p), out[1].ctypes.data_as(p),
                                   out[2].ctypes.data_as(p), self.n),
        )
        return tuple(a[: self.n] for a in out)


# ---- slab hash client (reference data_structs/chaining_table.cuh)
class SlabTable:
    """Chained slab hash table whose 120-byte slabs come from the allocator
    (gal_slab_* in include/gallatin_workloads.h)."""

    def __init__(self, allocator, buckets, seed=123456):
        self._w = _new_workload(allocator._h)
        h = ctypes.c_void_p()
        rc = _wl.gal_slab_create(self._w, buckets, seed, ctypes.byref(h))
        if rc != _capi.GAL_OK:
            _wl.gal_workload_destroy(self._w)
            _capi.check("gal_slab_create", rc)
        self._t = h
        self.allocator = allocator
        self.buckets = buckets
        allocator._adopt(self)

    def insert(self, n, key_seed):
        r = _capi.GalSlabResult()
        _capi.check("gal_slab_insert", _wl.gal_slab_insert(self._t, n, key_seed, ctypes.byref(r)))
        return r.as_dict()

    def query(self, n, key_seed):
        r = _capi.GalSlabResult()
        _capi.check("gal_slab_query", _wl.gal_slab_query(self._t, n, key_seed, ctypes.byref(r)))
        return r.as_dict()

    def destroy(self):
        freed = ctypes.c_uint64()
        if self._t:
            _capi.check("gal_slab_destroy", _wl.gal_slab_destroy(self._t, ctypes.byref(freed)))
            self._t = None
        if self._w:
            _capi.check("gal_workload_destroy", _wl.gal_workload_destroy(self._w))
            self._w = None
        return freed.value

    def _release(self):
        self.destroy()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


# ---- device atomic rates (the bound of the slice path; no reference counterpart)
ATOMIC_DISTINCT = 0
ATOMIC_WAVE_HOT = 1


def atomic_rate(mode, threads, iters, addresses=1):
    """Returned agent-scope 32-bit atomic adds per second (see gal_atomic_rate)."""
    out = ctypes.c_double()
    _capi.check("gal_atomic_rate", _wl.gal_atomic_rate(mode, threads, iters, addresses, ctypes.byref(out)))
    return out.value
