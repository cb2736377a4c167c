"""The C-ABI libraries load and export every entry point their headers
declare: libgallatin_amd.so (include/gallatin_amd.h) and
libgallatin_workloads.so (include/gallatin_workloads.h). Nothing else leaves
either library, and the workloads library uses the allocator library's one
mapping.

No compute calls: this runs without a GPU.
"""
import ctypes
import os
import re
import subprocess

from gallatin_amd import _capi


def declared_functions(header=_capi.HEADER_PATH):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|uint64_t)\s+(gal_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gal_create" in names and "gal_replay" in names and "gal_init_global_allocator" in names
    assert len(names) >= 19


def test_workloads_header_declares_entry_points():
    names = declared_functions(_capi.WL_HEADER_PATH)
    assert "gal_workload_create" in names and "gal_run_mixed" in names and "gal_serve_run" in names
    # no entry point is declared by both headers
    assert not set(names) & set(declared_functions())


def test_library_exports_every_declared_symbol():
    for path, header in ((_capi.LIB_PATH, _capi.HEADER_PATH), (_capi.WL_PATH, _capi.WL_HEADER_PATH)):
        lib = ctypes.CDLL(path)
        missing = [n for n in declared_functions(header) if not hasattr(lib, n)]
        assert missing == [], path


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_libraries_export_only_their_entry_points():
    # kernels and module-scope device variables stay inside each library, so
    # the two libraries' copies of the same templates do not bind to each other
    assert exported(_capi.LIB_PATH) == set(declared_functions())
    assert exported(_capi.WL_PATH) == set(declared_functions(_capi.WL_HEADER_PATH))


def test_allocator_library_mapped_once():
    maps = open("/proc/self/maps").read()
    paths = {line.split()[-1] for line in maps.splitlines() if line.endswith("libgallatin_amd.so")}
    assert paths == {os.path.realpath(_capi.LIB_PATH)}


def test_python_binding_covers_headers():
    assert set(declared_functions()) <= set(_capi._SIGNATURES)
    assert set(declared_functions(_capi.WL_HEADER_PATH)) <= set(_capi._WL_SIGNATURES)


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
