"""The C-ABI library loads and exports every entry point include/gallatin_amd.h declares.

No compute calls: this runs without a GPU.
"""
import ctypes
import re

from gallatin_amd import _capi


def declared_functions():
    text = open(_capi.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|uint64_t)\s+(gal_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gal_create" in names and "gal_replay" in names and "gal_init_global_allocator" in names
    assert len(names) >= 19


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_python_binding_covers_header():
    assert set(declared_functions()) <= set(_capi._SIGNATURES)


def test_supported_triples():
    import gallatin_amd

    triples = gallatin_amd.supported_triples()
    assert (16 << 20, 16, 4096) in triples
    for seg, mn, mx in triples:
        assert seg >= mx * 4096 and mn <= mx
        for v in (seg, mn, mx):
            assert v & (v - 1) == 0


def test_create_rejects_unsupported_triple():
    h = ctypes.c_void_p()
    rc = _capi.lib.gal_create(3 << 20, 16, 4096, 1 << 30, 1, ctypes.byref(h))
    assert rc == _capi.GAL_ERR_UNSUPPORTED
    assert _capi.lib.gal_destroy(None) == _capi.GAL_ERR_INVALID
