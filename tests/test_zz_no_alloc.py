"""The entry point that allocates nothing (gal_group_replay_noalloc) against
the one it varies (gal_group_replay), through the C-ABI.

Two allocators built alike run the same multi-lane op streams, one through
gal_group_replay (host buffers, library staging) and one through
gal_group_replay_noalloc (device buffers of the caller, the library's fixed
scratch array): results and whole metadata images must be equal bit for bit,
on two calls in a row, the second with fewer ops. Sizes above the scratch
limit are refused before anything runs, host pointers are refused, and an
invalid op found on the device stops the replay with the ops before it done.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import gallatin_amd as ga
from gallatin_amd import _capi
from gallatin_amd import snapshot as snap
from oracle import oracle
from tests import group_streams

pytestmark = pytest.mark.gpu

MiB = 1 << 20
NULL = (1 << 64) - 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class HipBuffers:
    """Device memory the caller owns, through the HIP runtime itself."""

    def __init__(self):
        self.hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.held = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.held.append(p.value)
        return p.value

    def upload(self, host_addr, nbytes):
        d = self.alloc(nbytes)
        assert self.hip.hipMemcpy(d, host_addr, nbytes, 1) == 0  # host to device
        return d

    def download(self, d, nbytes):
        out = np.zeros(nbytes, dtype=np.uint8)
        assert self.hip.hipMemcpy(out.ctypes.data, d, nbytes, 2) == 0  # device to host
        return out

    def free_all(self):
        for p in self.held:
            self.hip.hipFree(p)
        self.held = []


@pytest.fixture
def hipbuf():
    b = HipBuffers()
    yield b
    b.free_all()


def to_device(buf, ops):
    arr, n = ga.group_ops_array(ops)
    d_ops = buf.upload(ctypes.addressof(arr), ctypes.sizeof(_capi.GalGroupOp) * n)
    d_res = buf.alloc(n * 64 * 8)
    return d_ops, d_res, n


def results_of(buf, d_res, n):
    return buf.download(d_res, n * 64 * 8).view(np.uint64).reshape(n, 64)


def test_noalloc_equals_group_replay_on_two_calls(hipbuf):
    pool, seed = 1 << 30, 71
    o = oracle.Oracle(pool, seed)
    first, _, _ = group_streams.build(o, seed, 1500, [16, 32, 64, 256, 1000, 4096, 8192, 16384],
                                      (40000, MiB, 20 * MiB), large_p=0.05)
    second, _, _ = group_streams.build(o, seed + 1, 400, [16, 48, 128, 2048, 4096])
    assert len(second) < len(first)
    a = ga.Gallatin.generate_on_device(pool, seed)
    b = ga.Gallatin.generate_on_device(pool, seed)
    try:
        for k, ops in enumerate((first, second)):
            say(f"[noalloc] call {k}: {len(ops)} ops")
            expect = a.group_replay(ops)
            d_ops, d_res, n = to_device(hipbuf, ops)
            assert b.group_replay_noalloc(d_ops, n, d_res) == 0
            got = results_of(hipbuf, d_res, n)
            bad = np.nonzero((got != expect).any(axis=1))[0]
            assert len(bad) == 0, f"call {k} op {bad[0]}: {got[bad[0]].tolist()} vs {expect[bad[0]].tolist()}"
            sa, sb = a.snapshot(), b.snapshot()
            assert np.array_equal(sa, sb), snap.first_difference(sa, sb)
        # both calls together are the oracle's streams
        assert np.array_equal(b.snapshot(), o.snapshot())
        assert a.errors() == (0, 0) and b.errors() == (0, 0)
    finally:
        a.free_on_device()
        b.free_on_device()


def test_noalloc_refuses_what_it_cannot_hold(hipbuf):
    g = ga.Gallatin.generate_on_device(1 << 30, 3)
    try:
        before = g.snapshot()
        n = _capi.GAL_NOALLOC_MAX_OPS + 1
        d_ops = hipbuf.alloc(ctypes.sizeof(_capi.GalGroupOp) * n)
        d_res = hipbuf.alloc(n * 64 * 8)
        assert g.group_replay_noalloc(d_ops, n, d_res) == _capi.GAL_ERR_TOO_LARGE
        # host memory is not device memory
        host_ops, _ = ga.group_ops_array([(0, 0, 1, [16] * 64, [0] * 64)])
        host_res = (ctypes.c_uint64 * 64)()
        assert g.group_replay_noalloc(ctypes.addressof(host_ops), 1, ctypes.addressof(host_res)) == _capi.GAL_ERR_INVALID
        assert np.array_equal(g.snapshot(), before)
    finally:
        g.free_on_device()


def test_noalloc_stops_at_an_invalid_op(hipbuf):
    pool, seed = 1 << 30, 72
    lanes = (1 << 8) - 1
    sizes, tids = [64] * 64, list(range(64))
    free_lane0 = [0] * 64
    ops = [(0, 5, lanes, sizes, tids), (1, 0, 1, free_lane0, [0] * 64), (1, 0, 1, free_lane0, [0] * 64),
           (0, 5, lanes, sizes, tids)]
    o = oracle.Oracle(pool, seed)
    out = o.group_malloc(lanes, sizes, tids, 5)
    assert o.free(out[0]) == 0
    g = ga.Gallatin.generate_on_device(pool, seed)
    try:
        d_ops, d_res, n = to_device(hipbuf, ops)
        assert g.group_replay_noalloc(d_ops, n, d_res) == _capi.GAL_ERR_INVALID
        got = results_of(hipbuf, d_res, n)
        assert got[0][:8].tolist() == out[:8] and got[1][0] == 0
        assert np.array_equal(g.snapshot(), o.snapshot())
        # the failed call leaves nothing behind in the scratch: the next call works
        more = [(0, 9, lanes, sizes, tids)]
        expect = o.group_malloc(lanes, sizes, tids, 9)
        d_ops, d_res, n = to_device(hipbuf, more)
        assert g.group_replay_noalloc(d_ops, n, d_res) == 0
        assert results_of(hipbuf, d_res, n)[0][:8].tolist() == expect[:8]
        assert np.array_equal(g.snapshot(), o.snapshot())
        assert g.errors() == (0, 0)
    finally:
        g.free_on_device()
