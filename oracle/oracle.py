"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE, NOT PRODUCT CODE: only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg load this, and only as the checker or as the
timed CPU baseline.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the same sources (e.g. the sanitizer build)
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "build", "liboracle.so")
NULL = (1 << 64) - 1

u64 = ctypes.c_uint64
u32 = ctypes.c_uint32
vp = ctypes.c_void_p
pu64 = ctypes.POINTER(u64)


class _Op(ctypes.Structure):
    _fields_ = [("kind", u32), ("slot", u32), ("arg", u64), ("tid", u64)]


class _GroupOp(ctypes.Structure):
    _fields_ = [("kind", u32), ("slot", u32), ("lanes", u64), ("arg", u64 * 64), ("tid", u64 * 64)]


_SIG = {
    "orc_murmur64a": (u64, [u64, u64]),
    "orc_first_bit_bigger": (ctypes.c_int, [u64]),
    "orc_cfcll": (ctypes.c_int, [u64]),
    "orc_splitmix64": (u64, [u64]),
    "orc_reregister": (ctypes.c_int, [u32, u64]),
    "orc_route": (ctypes.c_int, [u64, u64, u64, u64, ctypes.POINTER(u32), ctypes.POINTER(u32), pu64]),
    "orc_veb_create": (vp, [u64, u64, ctypes.c_int]),
    "orc_veb_destroy": (None, [vp]),
    "orc_veb_layers": (ctypes.c_int, [vp]),
    "orc_veb_layer_words": (u64, [vp, ctypes.c_int]),
    "orc_veb_get_words": (None, [vp, ctypes.c_int, pu64]),
    "orc_veb_set_words": (None, [vp, ctypes.c_int, pu64]),
    "orc_veb_insert": (ctypes.c_int, [vp, u64]),
    "orc_veb_remove": (ctypes.c_int, [vp, u64]),
    "orc_veb_query": (ctypes.c_int, [vp, u64]),
    "orc_veb_successor": (u64, [vp, u64]),
    "orc_veb_lock_offset": (u64, [vp, u64]),
    "orc_veb_malloc_first": (u64, [vp]),
    "orc_veb_find_random": (u64, [vp, u64]),
    "orc_veb_gather_multiple": (u64, [vp, u64]),
    "orc_veb_return_multiple": (None, [vp, u64, u64]),
    "orc_create": (ctypes.c_int, [u64, u64, u64, u64, u64, ctypes.POINTER(vp)]),
    "orc_destroy": (None, [vp]),
    "orc_malloc": (u64, [vp, u64, u64, u32]),
    "orc_free": (ctypes.c_int, [vp, u64]),
    "orc_replay": (ctypes.c_int, [vp, ctypes.POINTER(_Op), u64, pu64]),
    "orc_group_replay": (ctypes.c_int, [vp, ctypes.POINTER(_GroupOp), u64, pu64]),
    "orc_group_malloc": (ctypes.c_int, [vp, u64, pu64, pu64, u32, ctypes.c_int, pu64]),
    "orc_group_free": (ctypes.c_int, [vp, u64, pu64]),
    "orc_snapshot": (ctypes.c_int, [vp, vp, u64, pu64]),
    "orc_faults": (u64, [vp]),
    "orc_bench_mixed": (ctypes.c_int, [u64, u64, u64, u32, pu64, pu64, ctypes.POINTER(ctypes.c_double)]),
    "orc_bench_sysmalloc_size": (
        ctypes.c_int,
        [u64, u32, u64, u64, pu64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u32)],
    ),
    "orc_bench_sysmalloc": (
        ctypes.c_int,
        [
            u64,
            u32,
            u64,
            pu64,
            ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(u32),
        ],
    ),
}


def load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIG.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


def murmur64a(key, seed):
    return lib.orc_murmur64a(key, seed)


def route(seg, mn, mx, size):
    p, t, c = u32(), u32(), u64()
    lib.orc_route(seg, mn, mx, size, ctypes.byref(p), ctypes.byref(t), ctypes.byref(c))
    return p.value, t.value, c.value


class Veb:
    def __init__(self, universe, seed=0, full=False):
        self.h = lib.orc_veb_create(universe, seed, int(full))
        self.universe = universe

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_veb_destroy(self.h)
            self.h = None

    def layers(self):
        return lib.orc_veb_layers(self.h)

    def words(self, layer):
        n = lib.orc_veb_layer_words(self.h, layer)
        out = np.zeros(n, dtype=np.uint64)
        lib.orc_veb_get_words(self.h, layer, out.ctypes.data_as(pu64))
        return out

    def set_words(self, layer, words):
        w = np.ascontiguousarray(words, dtype=np.uint64)
        lib.orc_veb_set_words(self.h, layer, w.ctypes.data_as(pu64))

    def insert(self, v):
        return bool(lib.orc_veb_insert(self.h, v))

    def remove(self, v):
        return bool(lib.orc_veb_remove(self.h, v))

    def query(self, v):
        return bool(lib.orc_veb_query(self.h, v))

    def successor(self, v):
        return lib.orc_veb_successor(self.h, v)

    def lock_offset(self, v):
        return lib.orc_veb_lock_offset(self.h, v)

    def malloc_first(self):
        return lib.orc_veb_malloc_first(self.h)

    def find_random(self, tid):
        return lib.orc_veb_find_random(self.h, tid)

    def gather_multiple(self, n):
        return lib.orc_veb_gather_multiple(self.h, n)

    def return_multiple(self, start, n):
        lib.orc_veb_return_multiple(self.h, start, n)


class Oracle:
    """Sequential CPU model of Gallatin<segment, min, max>."""

    def __init__(self, pool_bytes, seed, segment=16 << 20, min_size=16, max_size=4096):
        h = vp()
        rc = lib.orc_create(segment, min_size, max_size, pool_bytes, seed, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"orc_create failed: {rc}")
        self.h = h
        self.segment, self.min_size, self.max_size = segment, min_size, max_size

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_destroy(self.h)
            self.h = None

    def malloc(self, size, tid=0, slot=0):
        return lib.orc_malloc(self.h, size, tid, slot)

    def free(self, offset):
        return lib.orc_free(self.h, offset)

    def faults(self):
        return lib.orc_faults(self.h)

    def replay(self, ops):
        ops = list(ops)
        arr = (_Op * max(1, len(ops)))()
        for i, (kind, arg, tid, slot) in enumerate(ops):
            arr[i].kind, arr[i].arg, arr[i].tid, arr[i].slot = kind, arg, tid, slot
        out = (u64 * max(1, len(ops)))()
        rc = lib.orc_replay(self.h, arr, len(ops), out)
        if rc != 0:
            raise RuntimeError(f"orc_replay failed: {rc}")
        return np.frombuffer(out, dtype=np.uint64, count=len(ops)).copy()

    def group_replay(self, ops):
        """ops: (kind, slot, lanes, args, tids) per op; returns an (n, 64) array."""
        ops = list(ops)
        arr = (_GroupOp * max(1, len(ops)))()
        for i, (kind, slot, lanes, args, tids) in enumerate(ops):
            arr[i].kind, arr[i].slot, arr[i].lanes = kind, slot, lanes
            for lane in range(64):
                if (lanes >> lane) & 1:
                    arr[i].arg[lane] = int(args[lane])
                    arr[i].tid[lane] = int(tids[lane])
        out = (u64 * max(1, 64 * len(ops)))()
        rc = lib.orc_group_replay(self.h, arr, len(ops), out)
        if rc != 0:
            raise RuntimeError(f"orc_group_replay failed: {rc}")
        return np.frombuffer(out, dtype=np.uint64, count=64 * len(ops)).copy().reshape(len(ops), 64)

    def group_malloc(self, lanes, sizes, tids, slot, spread=False):
        """malloc by the lanes of `lanes` together; sizes/tids: 64 entries by
        lane. spread: the claims malloc makes under load (GAL_OP_MALLOC_SPREAD)."""
        sz = (u64 * 64)(*[int(x) for x in sizes])
        td = (u64 * 64)(*[int(x) for x in tids])
        out = (u64 * 64)()
        rc = lib.orc_group_malloc(self.h, lanes, sz, td, slot, int(bool(spread)), out)
        if rc != 0:
            raise RuntimeError(f"orc_group_malloc failed: {rc}")
        return list(out)

    def group_free(self, lanes, offsets):
        """free by the lanes of `lanes` together; offsets: 64 entries by lane."""
        off = (u64 * 64)(*[int(x) for x in offsets])
        return lib.orc_group_free(self.h, lanes, off)

    def snapshot(self):
        need = u64()
        lib.orc_snapshot(self.h, None, 0, ctypes.byref(need))
        buf = np.zeros(need.value, dtype=np.uint8)
        rc = lib.orc_snapshot(self.h, buf.ctypes.data_as(vp), need.value, None)
        if rc != 0:
            raise RuntimeError(f"orc_snapshot failed: {rc}")
        return buf


def bench_mixed(pool_bytes, seed, threads_per_round, rounds):
    ops, miss, sec = u64(), u64(), ctypes.c_double()
    rc = lib.orc_bench_mixed(pool_bytes, seed, threads_per_round, rounds, ctypes.byref(ops), ctypes.byref(miss),
                             ctypes.byref(sec))
    if rc != 0:
        raise RuntimeError(f"orc_bench_mixed failed: {rc}")
    return ops.value, miss.value, sec.value


def bench_sysmalloc(ops_per_thread, threads=0, seed=1, size=0):
    """glibc malloc/free on `threads` host threads (0: all); size 0 draws the
    log-uniform 16 B-4 KiB law, otherwise every request is `size` bytes."""
    ops, sec, p50, p99, used = u64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), u32()
    rc = lib.orc_bench_sysmalloc_size(ops_per_thread, threads, seed, size, ctypes.byref(ops), ctypes.byref(sec),
                                      ctypes.byref(p50), ctypes.byref(p99), ctypes.byref(used))
    return {"ops": ops.value, "seconds": sec.value, "p50_ns": p50.value, "p99_ns": p99.value,
            "threads": used.value, "rc": rc}
