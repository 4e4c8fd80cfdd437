"""CPU-side checks of the C ABI: the HIP library loads, exports every entry
point include/gzgpu.h declares, and refuses to run without a gfx950 GPU
(no silent CPU fallback)."""
import ctypes
import os
import re

import pytest

from gzippy_amd import _native, build

HEADER = os.path.join(build.INCLUDE, "gzgpu.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gzgpu_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(build.gpu_lib_path()):
        build.build_gpu()
    return _native.load_library()


def test_header_declares_the_binding(lib):
    decl = _declared()
    assert decl == sorted(_native.EXPORTS)
    for name in decl:
        assert hasattr(lib, name), name


def test_exports_in_dynamic_symbol_table(lib):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", build.gpu_lib_path()], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in _declared():
        assert name in syms, name


def test_status_strings(lib):
    # ParallelError Display texts (single_member.rs:1415-1437)
    assert lib.gzgpu_status_string(1) == b"invalid gzip header"
    assert lib.gzgpu_status_string(4) == b"output size mismatch"
    assert lib.gzgpu_status_string(5) == b"CRC32 mismatch"
    assert lib.gzgpu_status_string(3) == b"chunk decode failed"


def test_no_gpu_means_error_not_fallback(lib):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible; covered by the gpu tests")
    h = ctypes.c_void_p()
    assert lib.gzgpu_ctx_create(0, ctypes.byref(h)) == _native.GPU_UNAVAILABLE
    with pytest.raises(_native.GzgpuError):
        _native.Context(0)
