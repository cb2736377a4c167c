"""Host logic of handle ownership (no GPU): a handle that batches, serving
loops or slab tables still use is not destroyed.

The library refuses gal_destroy with GAL_ERR_STATE while such objects exist
(they hold the handle's stream and allocator). The Python wrapper releases
the objects it made on a handle before destroying it. Checked here against a
stand-in for the C library that keeps the same child count and refuses the
same way.
"""
import ctypes

import pytest

import gallatin_amd as ga
from gallatin_amd import _capi


class FakeLib:
    """The C-ABI's ownership rule: destroy refused while children exist."""

    def __init__(self):
        self.children = 0
        self.destroyed = False
        self.calls = []

    def _make(self, out):
        self.children += 1
        out._obj.value = 1000 + self.children
        return _capi.GAL_OK

    def gal_batch_create(self, h, capacity, out):
        self.calls.append("batch_create")
        return self._make(out)

    def gal_serve_create(self, h, capacity, out):
        self.calls.append("serve_create")
        return self._make(out)

    def gal_slab_create(self, h, buckets, seed, out):
        self.calls.append("slab_create")
        return self._make(out)

    def gal_batch_destroy(self, b):
        self.calls.append("batch_destroy")
        self.children -= 1
        return _capi.GAL_OK

    def gal_serve_destroy(self, s):
        self.calls.append("serve_destroy")
        self.children -= 1
        return _capi.GAL_OK

    def gal_slab_destroy(self, t, freed):
        self.calls.append("slab_destroy")
        self.children -= 1
        return _capi.GAL_OK

    def gal_destroy(self, h):
        self.calls.append("destroy")
        if self.children:
            return _capi.GAL_ERR_STATE
        self.destroyed = True
        return _capi.GAL_OK


@pytest.fixture
def fake(monkeypatch):
    lib = FakeLib()
    monkeypatch.setattr(ga, "_lib", lib)
    return lib


def test_children_are_released_before_the_handle(fake):
    g = ga.Gallatin(ctypes.c_void_p(1), None)
    b = ga.Batch(g, 16)
    s = ga.Serving(g, 16)
    t = ga.SlabTable(g, 16)
    assert fake.children == 3
    g.free_on_device()
    assert fake.destroyed and fake.children == 0
    assert fake.calls[-1] == "destroy"
    assert b._b is None and s._s is None and t._t is None


def test_a_closed_child_is_not_released_twice(fake):
    g = ga.Gallatin(ctypes.c_void_p(1), None)
    b = ga.Batch(g, 16)
    b.close()
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
