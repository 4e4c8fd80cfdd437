"""gzgpu_decompress_device_noalloc: the single-member device decode with
every buffer from the library's static scratch array.  Checked through the C
ABI against gzgpu_decompress_device on the same member, byte for byte with
CRC32 and length, on two calls in a row (the second smaller, so it runs on
scratch that still holds the first call's contents), and the limits."""
import ctypes
import zlib

import numpy as np
import pytest

from gzippy_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _call(ctx, fn, d_in, d_out):
    n = ctypes.c_size_t()
    crc = ctypes.c_uint32()
    st = fn(ctx.h, ctypes.c_void_p(d_in.data_ptr()), d_in.numel(), ctypes.c_void_p(d_out.data_ptr()),
            d_out.numel(), ctypes.byref(n), ctypes.byref(crc), None)
    return st, n.value, crc.value


@pytest.mark.parametrize("kinds", [("text", "u32"), ("runs", "random"), ("u32", "text")])
def test_noalloc_equals_device_decode(gpu, kinds):
    import torch
    ctx = gpu.context()
    lib = ctx.lib
    torch.cuda.synchronize()
    sizes = (24 << 20, (3 << 20) + 12345)  # second call smaller
    for i, (kind, size) in enumerate(zip(kinds, sizes)):
        raw = W.generate(kind, size, seed=900 + i)
        gz = W.gzip_compress(raw, (6, 1)[i])
        say("noalloc", kind, size, gz.size)
        d_in = torch.from_numpy(gz).cuda()
        ref = torch.zeros(raw.size, dtype=torch.uint8, device="cuda")
        out = torch.full((raw.size,), 0xAB, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        st_r, n_r, crc_r = _call(ctx, lib.gzgpu_decompress_device, d_in, ref)
        assert st_r == 0, ctx.error(st_r)
        st, n, crc = _call(ctx, lib.gzgpu_decompress_device_noalloc, d_in, out)
        assert st == 0, ctx.error(st)
        assert (n, crc) == (n_r, crc_r) == (raw.size, zlib.crc32(raw.tobytes()))
        assert torch.equal(out, ref)
        assert np.array_equal(out.cpu().numpy(), raw)


def test_noalloc_limits(gpu):
    """Inputs past the limits are refused before any device work, with a
    message; the context keeps working afterwards."""
    import torch
    ctx = gpu.context()
    lib = ctx.lib
    big = torch.zeros((32 << 20) + 1, dtype=torch.uint8, device="cuda")
    out = torch.zeros(16, dtype=torch.uint8, device="cuda")
    st, _, _ = _call(ctx, lib.gzgpu_decompress_device_noalloc, big, out)
    assert st == 10 and "limit" in str(ctx.error(st))  # GZGPU_INVALID_ARGUMENT
    # a member whose trailer says 64 MiB
    raw = W.generate("runs", 64 << 20, seed=5)
    gz = W.gzip_compress(raw, 9)
    d_in = torch.from_numpy(gz).cuda()
    out = torch.zeros(raw.size, dtype=torch.uint8, device="cuda")
    st, _, _ = _call(ctx, lib.gzgpu_decompress_device_noalloc, d_in, out)
    assert st == 10 and "ISIZE" in str(ctx.error(st))
    # and then a good one
    raw = W.generate("text", 1 << 20, seed=6)
    gz = W.gzip_compress(raw, 6)
    d_in = torch.from_numpy(gz).cuda()
    out = torch.zeros(raw.size, dtype=torch.uint8, device="cuda")
    st, n, crc = _call(ctx, lib.gzgpu_decompress_device_noalloc, d_in, out)
    assert st == 0 and n == raw.size and np.array_equal(out.cpu().numpy(), raw)
