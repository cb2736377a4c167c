"""ctypes bindings of libgallatin_amd.so (the allocator's C-ABI,
include/gallatin_amd.h) and libgallatin_workloads.so (the benchmark and test
workloads on top of it, include/gallatin_workloads.h).

The libraries are the product: there is no CPU fallback. If they have not
been built, importing this module raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgallatin_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "gallatin_amd.h")
WL_PATH = os.path.join(_HERE, "lib", "libgallatin_workloads.so")
WL_HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "gallatin_workloads.h")

GAL_OK = 0
GAL_ERR_INVALID = 1
GAL_ERR_UNSUPPORTED = 2
GAL_ERR_OOM = 3
GAL_ERR_HIP = 4
GAL_ERR_BUFFER = 5
GAL_ERR_STATE = 6
GAL_ERR_TOO_LARGE = 7
ERROR_NAMES = {
    GAL_OK: "ok",
    GAL_ERR_INVALID: "invalid argument",
    GAL_ERR_UNSUPPORTED: "unsupported (segment, min, max) triple",
    GAL_ERR_OOM: "out of device memory",
    GAL_ERR_HIP: "HIP runtime error",
    GAL_ERR_BUFFER: "buffer too small",
    GAL_ERR_STATE: "allocator state error (or the handle is still in use)",
    GAL_ERR_TOO_LARGE: "request above the entry point's fixed limit",
}

GAL_OP_MALLOC = 0
GAL_OP_FREE = 1
GAL_NULL_OFFSET = (1 << 64) - 1
GAL_VERIFY_STORE = 0
GAL_VERIFY_ATOMIC = 1
GAL_FORM_PER_THREAD = 0
GAL_FORM_BATCHED = 1
GAL_FORM_GRAPH = 2
GAL_NOALLOC_MAX_OPS = 8192

u64 = ctypes.c_uint64
u32 = ctypes.c_uint32
vp = ctypes.c_void_p


class GalOp(ctypes.Structure):
    _fields_ = [("kind", u32), ("slot", u32), ("arg", u64), ("tid", u64)]


class GalGroupOp(ctypes.Structure):
    _fields_ = [("kind", u32), ("slot", u32), ("lanes", u64), ("arg", u64 * 64), ("tid", u64 * 64)]


class GalInfo(ctypes.Structure):
    _fields_ = [
        ("segment_size", u64),
        ("min_size", u64),
        ("max_size", u64),
        ("num_segments", u64),
        ("blocks_per_segment", u64),
        ("num_trees", u64),
        ("veb_layers", u64),
        ("total_slots", u64),
        ("pool_device_ptr", u64),
        ("device_handle", u64),
        ("snapshot_bytes", u64),
        ("metadata_bytes", u64),
    ]


class GalRunResult(ctypes.Structure):
    _fields_ = [
        ("mallocs", u64),
        ("frees", u64),
        ("misses", u64),
        ("tag_errors", u64),
        ("alloc_errors", u64),
        ("launches", u64),
        ("malloc_kernel_ms", ctypes.c_double),
        ("free_kernel_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class GalRequest(ctypes.Structure):
    _fields_ = [("arrival_ns", u64), ("id", u64), ("pages", u32), ("page_bytes", u32)]


class GalServeWorkload(ctypes.Structure):
    _fields_ = [
        ("requests", u64),
        ("rate_per_s", ctypes.c_double),
        ("median_pages", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("max_pages", u32),
        ("page_bytes", u32),
        ("size_classes", u32),
        ("unused", u32),
        ("seed", u64),
    ]


class GalServeResult(ctypes.Structure):
    _fields_ = [
        ("requests_done", u64),
        ("requests_aborted", u64),
        ("page_mallocs", u64),
        ("page_frees", u64),
        ("page_retries", u64),
        ("admission_refusals", u64),
        ("peak_active", u64),
        ("peak_committed", u64),
        ("alloc_errors", u64),
        ("slots", u64),
        ("run", u64),
        ("headroom", u64),
        ("kernel_ms", ctypes.c_double),
        ("makespan_ms", ctypes.c_double),
        ("page_malloc_p50_ns", ctypes.c_double),
        ("page_malloc_p99_ns", ctypes.c_double),
        ("page_malloc_max_ns", ctypes.c_double),
        ("queue_p50_ns", ctypes.c_double),
        ("queue_p99_ns", ctypes.c_double),
        ("latency_p50_ns", ctypes.c_double),
        ("latency_p99_ns", ctypes.c_double),
        ("window_pages", u64),
        ("window_ms", ctypes.c_double),
        ("first_dry_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class GalSlabResult(ctypes.Structure):
    _fields_ = [
        ("keys", u64),
        ("failed", u64),
        ("wrong", u64),
        ("slabs", u64),
        ("max_chain", u64),
        ("alloc_errors", u64),
        ("kernel_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class GalFillReport(ctypes.Structure):
    _fields_ = [
        ("num_trees", u64),
        ("free_segments", u64),
        ("max_segments", u64),
        ("listed_segments", u64 * 32),
        ("live_segments", u64 * 32),
        ("blocks_available", u64 * 32),
        ("blocks_in_use", u64 * 32),
        ("overhead_bytes", u64),
        ("empty_pinned_slots", u64 * 32),
    ]


GAL_SERVE_DEFAULT_HEADROOM = (1 << 64) - 1

_SIGNATURES = {
    "gal_supported": (ctypes.c_int, [ctypes.POINTER(u64), ctypes.c_int]),
    "gal_create": (ctypes.c_int, [u64, u64, u64, u64, u64, ctypes.POINTER(vp)]),
    "gal_destroy": (ctypes.c_int, [vp]),
    "gal_get_info": (ctypes.c_int, [vp, ctypes.POINTER(GalInfo)]),
    "gal_print_stats": (ctypes.c_int, [vp]),
    "gal_fill": (ctypes.c_int, [vp, ctypes.POINTER(GalFillReport)]),
    "gal_snapshot": (ctypes.c_int, [vp, vp, u64, ctypes.POINTER(u64)]),
    "gal_replay": (ctypes.c_int, [vp, ctypes.POINTER(GalOp), u64, ctypes.POINTER(u64)]),
    "gal_group_replay": (ctypes.c_int, [vp, ctypes.POINTER(GalGroupOp), u64, ctypes.POINTER(u64)]),
    "gal_group_replay_noalloc": (ctypes.c_int, [vp, vp, u64, vp]),
    "gal_errors": (ctypes.c_int, [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
    "gal_error_sites": (ctypes.c_int, [vp, ctypes.POINTER(u32), u32]),
    "gal_stats": (ctypes.c_int, [vp, ctypes.POINTER(u32), u32]),
    "gal_init_global_allocator": (ctypes.c_int, [u64, u64, ctypes.c_int]),
    "gal_free_global_allocator": (ctypes.c_int, []),
    "gal_print_global_stats": (ctypes.c_int, []),
    "gal_global_replay": (ctypes.c_int, [ctypes.POINTER(GalOp), u64, ctypes.POINTER(u64), ctypes.POINTER(u32)]),
    "gal_create_host": (ctypes.c_int, [u64, u64, u64, u64, u64, ctypes.POINTER(vp)]),
    "gal_allocated_size": (ctypes.c_int, [u64, u64, u64, u64, ctypes.POINTER(u64)]),
    "gal_init_global_allocator_host": (ctypes.c_int, [u64, u64, ctypes.c_int]),
    "gal_free_global_allocator_host": (ctypes.c_int, []),
    "gal_print_global_stats_host": (ctypes.c_int, []),
    "gal_init_global_allocator_combined": (ctypes.c_int, [u64, u64, u64, ctypes.c_int]),
    "gal_free_global_allocator_combined": (ctypes.c_int, []),
    "gal_print_global_stats_combined": (ctypes.c_int, []),
    "gal_global_handles": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.POINTER(vp)]),
    "gal_read_pool": (ctypes.c_int, [vp, u64, vp, u64]),
    "gal_free_bytes": (ctypes.c_int, [vp, ctypes.POINTER(u64)]),
    "gal_gather_limit": (ctypes.c_int, [ctypes.POINTER(u64), u64, ctypes.POINTER(u64)]),
    "gal_largest_run": (ctypes.c_int, [vp, ctypes.POINTER(u64)]),
    "gal_stream": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
    "gal_retain": (ctypes.c_int, [vp]),
    "gal_release": (ctypes.c_int, [vp]),
}

_WL_SIGNATURES = {
    "gal_workload_create": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
    "gal_workload_destroy": (ctypes.c_int, [vp]),
    "gal_global_smoke": (ctypes.c_int, [u64, u64, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
    "gal_global_plain": (ctypes.c_int, [u64, u64, u64, u32, ctypes.POINTER(GalRunResult)]),
    "gal_global_fused": (
        ctypes.c_int,
        [u64, u64, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64)],
    ),
    "gal_global_poison": (ctypes.c_int, [u64, ctypes.POINTER(ctypes.c_int32), u32]),
    "gal_run_plain": (ctypes.c_int, [vp, u64, u64, u64, u32, ctypes.POINTER(GalRunResult)]),
    "gal_mixed_reset": (ctypes.c_int, [vp, u64, u32, u64]),
    "gal_mixed_round": (ctypes.c_int, [vp, u32, u32, ctypes.POINTER(GalRunResult)]),
    "gal_mixed_drain": (ctypes.c_int, [vp, u32, ctypes.POINTER(GalRunResult)]),
    "gal_mixed_live": (
        ctypes.c_int,
        [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64), u64, ctypes.POINTER(u64)],
    ),
    "gal_run_mixed": (ctypes.c_int, [vp, u64, u32, u64, u32, ctypes.POINTER(GalRunResult)]),
    "gal_atomic_rate": (ctypes.c_int, [u32, u64, u32, u32, ctypes.POINTER(ctypes.c_double)]),
    "gal_batch_create": (ctypes.c_int, [vp, u64, ctypes.POINTER(vp)]),
    "gal_batch_destroy": (ctypes.c_int, [vp]),
    "gal_batch_generate": (ctypes.c_int, [vp, u64, u64, u32]),
    "gal_batch_set_sizes": (ctypes.c_int, [vp, ctypes.POINTER(u64), u64, u64]),
    "gal_batch_run": (ctypes.c_int, [vp, u32, u32, u32, u32, u32, ctypes.POINTER(GalRunResult)]),
    "gal_batch_malloc": (ctypes.c_int, [vp, u32, u32, ctypes.POINTER(GalRunResult)]),
    "gal_batch_free": (ctypes.c_int, [vp, ctypes.POINTER(GalRunResult)]),
    "gal_batch_offsets": (ctypes.c_int, [vp, ctypes.POINTER(u64), u64]),
    "gal_batch_order": (ctypes.c_int, [vp, ctypes.POINTER(u32), u64]),
    "gal_latency_samples": (
        ctypes.c_int,
        [vp, ctypes.POINTER(u32), u64, ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_double)],
    ),
    "gal_latency_keep": (ctypes.c_int, [vp, u32]),
    "gal_slab_create": (ctypes.c_int, [vp, u64, u64, ctypes.POINTER(vp)]),
    "gal_slab_insert": (ctypes.c_int, [vp, u64, u64, ctypes.POINTER(GalSlabResult)]),
    "gal_slab_query": (ctypes.c_int, [vp, u64, u64, ctypes.POINTER(GalSlabResult)]),
    "gal_slab_destroy": (ctypes.c_int, [vp, ctypes.POINTER(u64)]),
    "gal_churn": (ctypes.c_int, [vp, ctypes.POINTER(GalOp), u64, u64, ctypes.POINTER(u64), ctypes.POINTER(GalRunResult)]),
    "gal_serve_generate": (ctypes.c_int, [ctypes.POINTER(GalServeWorkload), vp]),
    "gal_serve_create": (ctypes.c_int, [vp, u64, ctypes.POINTER(vp)]),
    "gal_serve_destroy": (ctypes.c_int, [vp]),
    "gal_serve_load": (ctypes.c_int, [vp, vp, u64]),
    "gal_serve_run": (
        ctypes.c_int,
        [vp, u32, u32, u64, ctypes.c_double, ctypes.c_double, ctypes.POINTER(GalServeResult)],
    ),
    "gal_serve_results": (
        ctypes.c_int,
        [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64), u64],
    ),
}


class GallatinError(RuntimeError):
    def __init__(self, func, code):
        super().__init__(f"{func} failed: {ERROR_NAMES.get(code, code)} ({code})")
        self.code = code


def load(path=LIB_PATH, signatures=None):
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C gallatin_amd/csrc`"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in (_SIGNATURES if signatures is None else signatures).items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def check(func, code):
    if code != GAL_OK:
        raise GallatinError(func, code)


lib = load()
# after the allocator library, which it links against (the loader maps that once)
wl = load(WL_PATH, _WL_SIGNATURES)
