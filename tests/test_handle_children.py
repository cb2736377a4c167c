"""Host logic of handle ownership (no GPU): a handle that workload objects,
batches, serving loops or slab tables still use is not destroyed.

The allocator library refuses gal_destroy (and the global free calls) with
GAL_ERR_STATE while holds are taken on the handle (gal_retain); a workload
object of the workloads library holds its handle, and refuses its own
destroy while batches, serving loops or slab tables made on it exist. The
Python wrapper releases what it made on a handle before destroying it, and
drops the idle workload objects of global-allocator views before a global
free. Checked here against a stand-in for both libraries that keeps the same
counts and refuses the same way.
"""
import ctypes

import pytest

import gallatin_amd as ga
from gallatin_amd import _capi


class FakeLibs:
    """The two C-ABIs' ownership rules on one allocator handle."""

    def __init__(self):
        self.holds = 0  # holds on the handle
        self.children = {}  # workload id -> live batches / serving loops / slab tables
        self.owner = {}  # child id -> workload id
        self.next = 1000
        self.destroyed = False
        self.global_freed = False
        self.calls = []

    def _new(self, out):
        self.next += 1
        out._obj.value = self.next
        return self.next

    # workloads library
    def gal_workload_create(self, h, out):
        self.calls.append("workload_create")
        self.children[self._new(out)] = 0
        self.holds += 1
        return _capi.GAL_OK

    def gal_workload_destroy(self, w):
        self.calls.append("workload_destroy")
        if self.children[w.value]:
            return _capi.GAL_ERR_STATE
        del self.children[w.value]
        self.holds -= 1
        return _capi.GAL_OK

    def _child(self, what, w, out):
        self.calls.append(what)
        self.children[w.value] += 1
        self.owner[self._new(out)] = w.value
        return _capi.GAL_OK

    def _end(self, what, c):
        self.calls.append(what)
        self.children[self.owner.pop(c.value)] -= 1
        return _capi.GAL_OK

    def gal_batch_create(self, w, capacity, out):
        return self._child("batch_create", w, out)

    def gal_serve_create(self, w, capacity, out):
        return self._child("serve_create", w, out)

    def gal_slab_create(self, w, buckets, seed, out):
        return self._child("slab_create", w, out)

    def gal_batch_destroy(self, b):
        return self._end("batch_destroy", b)

    def gal_serve_destroy(self, s):
        return self._end("serve_destroy", s)

    def gal_slab_destroy(self, t, freed):
        return self._end("slab_destroy", t)

    def gal_run_plain(self, w, n, size, seed, verify, out):
        self.calls.append("run_plain")
        return _capi.GAL_OK

    # allocator library
    def gal_destroy(self, h):
        self.calls.append("destroy")
        if self.holds:
            return _capi.GAL_ERR_STATE
        self.destroyed = True
        return _capi.GAL_OK

    def gal_global_handles(self, device, host):
        device._obj.value = 1
        host._obj.value = None
        return _capi.GAL_OK

    def gal_get_info(self, h, info):
        return _capi.GAL_OK

    def gal_free_global_allocator(self):
        self.calls.append("free_global")
        if self.holds:
            return _capi.GAL_ERR_STATE
        self.global_freed = True
        return _capi.GAL_OK


@pytest.fixture
def fake(monkeypatch):
    libs = FakeLibs()
    monkeypatch.setattr(ga, "_lib", libs)
    monkeypatch.setattr(ga, "_wl", libs)
    return libs


def test_children_are_released_before_the_handle(fake):
    g = ga.Gallatin(ctypes.c_void_p(1), None)
    b = ga.Batch(g, 16)
    s = ga.Serving(g, 16)
    t = ga.SlabTable(g, 16)
    g.run_plain(64, 16, 1)  # the handle's own workload object
    assert fake.holds == 4 and sum(fake.children.values()) == 3
    g.free_on_device()
    assert fake.destroyed and fake.holds == 0 and not fake.children
    assert fake.calls[-1] == "destroy"
    assert b._b is None and s._s is None and t._t is None
    assert b._w is None and s._w is None and t._w is None and g._w is None


def test_a_closed_child_is_not_released_twice(fake):
    g = ga.Gallatin(ctypes.c_void_p(1), None)
    b = ga.Batch(g, 16)
    b.close()
    assert fake.holds == 0  # the batch's workload object went with it
    g.free_on_device()
    assert fake.calls.count("batch_destroy") == 1 and fake.destroyed


def test_the_library_refusal_surfaces_as_an_error(fake):
    g = ga.Gallatin(ctypes.c_void_p(1), None)
    ga.Batch(g, 16)
    g._children = []  # a child the wrapper does not know about
    with pytest.raises(ga.GallatinError) as e:
        g.free_on_device()
    assert e.value.code == _capi.GAL_ERR_STATE
    g._h = None  # the stand-in handle must not reach the real library later


def test_a_view_workload_does_not_block_the_global_free(fake):
    dev, host = ga.global_handles()
    assert host is None
    dev.run_plain(64, 16, 1)
    assert fake.holds == 1
    ga.free_global_allocator()
    assert fake.global_freed and fake.holds == 0 and dev._w is None


def test_a_live_table_on_a_view_still_blocks_the_global_free(fake):
    dev, _ = ga.global_handles()
    t = ga.SlabTable(dev, 16)
    with pytest.raises(ga.GallatinError) as e:
        ga.free_global_allocator()
    assert e.value.code == _capi.GAL_ERR_STATE and not fake.global_freed
    t.destroy()
    ga.free_global_allocator()
    assert fake.global_freed and fake.holds == 0
