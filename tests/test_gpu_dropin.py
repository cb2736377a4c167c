"""A user program built against the headers alone runs on the GPU.

gallatin_amd/examples/dropin: two translation units compiled with hipcc
-fgpu-rdc and only gallatin_amd/include on the include path, using the
reference's include line (<gallatin/allocators/global_allocator.cuh>) and
API (init_global_allocator, global_malloc/global_free called from device
functions of the other unit, print_global_stats, free_global_allocator, and
Gallatin<16 MiB, 16, 4096>::generate_on_device / malloc / free /
free_on_device), reference README.md:15-56. build() compiles it into
gallatin_amd/lib/dropin_example.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gallatin_amd", "lib", "dropin_example")


@pytest.mark.gpu
def test_dropin_program_runs():
    assert os.path.exists(EXE), "build() compiles the drop-in example"
    print("[dropin] running", EXE, flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "dropin ok" in r.stdout
