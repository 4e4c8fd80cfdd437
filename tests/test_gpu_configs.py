"""The full-size single-member workloads, decoded on the GPU and compared byte
for byte (and CRC32 / ISIZE) with the generator's bytes:

* 1 GiB of the text generator at level 6, one member (libdeflate);
* the mixed-entropy member: 64 MiB segments cycling text / uniform random /
  byte runs / u32 deltas, at level 1 and at level 9 (pigz-style zlib:
  segments encoded in parallel, one member), here 33 segments so that
  output offsets pass 2^31.

The 16 MiB text member at level 6 is tests/test_gpu_decompress.py::
test_text_member_16mib."""
import numpy as np
import pytest

from gzippy_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _decode_on_device(gpu, raw, gz):
    import torch
    d_in = torch.from_numpy(gz).cuda()
    d_raw = torch.from_numpy(raw).cuda()
    d_out = torch.zeros(raw.size, dtype=torch.uint8, device="cuda")
    n, crc, st = gpu.decompress_device(d_in, d_out)
    torch.cuda.synchronize()
    ok = n == raw.size and bool(torch.equal(d_out, d_raw))
    del d_in, d_raw, d_out
    torch.cuda.empty_cache()
    return ok, n, crc, st


def test_single_member_1gib_text_l6(gpu):
    raw, gz = W.text_member(1 << 30, seed=0x675A0001, level=6)
    say("1 GiB text L6", raw.size, "->", gz.size)
    ok, n, crc, st = _decode_on_device(gpu, raw, gz)
    say("stats", {k: st[k] for k in ("pieces", "redecoded", "inflate_launches", "ms_total")})
    assert ok, (n, raw.size)
    assert crc == W.trailer(gz)[0]


@pytest.fixture(scope="module")
def mixed_raw():
    n = 33 * W.SEG + 12345  # past 2^31 output bytes
    raw = W.generate_mixed(n, seed=0x675A0002, threads=16)
    yield raw


@pytest.mark.parametrize("level", [1, 9])
def test_mixed_member_past_2gib(gpu, mixed_raw, level):
    raw = mixed_raw
    gz = W.gzip_compress_segments(raw, level)
    say("mixed member L%d" % level, raw.size, "->", gz.size)
    assert raw.size > (1 << 31)
    ok, n, crc, st = _decode_on_device(gpu, raw, gz)
    say("stats", {k: st[k] for k in ("pieces", "redecoded", "inflate_launches", "ms_total", "ms_find",
                                     "ms_inflate", "ms_resolve")})
    assert ok, (n, raw.size)
    assert crc == W.trailer(gz)[0]
