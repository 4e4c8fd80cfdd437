"""Seekable index (src/decompress/index.rs): the index the GPU builds is in
index.rs's version-1 layout; every checkpoint is pinned independently by the
oracle (a decode from its bit offset with its window reproduces the
generator's bytes at its output offset); seeks return the generator's bytes
for any range; an index of another input is refused."""
import ctypes
import struct

import numpy as np
import pytest

from gzippy_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def parse_index(ix):
    assert ix[:6] == b"GZIDX\x01"
    deflate_off, total, npts = struct.unpack_from("<IQI", ix, 6)
    pts = []
    for i in range(npts):
        bit, out = struct.unpack_from("<QQ", ix, 24 + i * (16 + 32768))
        win = ix[24 + i * (16 + 32768) + 16: 24 + (i + 1) * (16 + 32768)]
        pts.append((bit, out, win))
    assert len(ix) == 24 + npts * (16 + 32768)
    return deflate_off, total, pts


@pytest.mark.parametrize("kind,level", [("text", 6), ("u32", 1), ("random", 6), ("runs", 9)])
def test_index_points_and_seeks(gpu, oracle, kind, level):
    raw = W.generate(kind, (48 << 20) + 123, seed=31 + level)
    gz = W.gzip_compress(raw, level)
    ix = gpu.build_index(gz, 1 << 20)
    deflate_off, total, pts = parse_index(ix)
    say("index", kind, level, len(pts), "points")
    assert total == raw.size and pts[0][0] == 0 and pts[0][1] == 0
    assert len(pts) >= 8
    outs = [p[1] for p in pts]
    assert all(b - a >= 1 << 20 for a, b in zip(outs, outs[1:]))
    rng = np.random.default_rng(7)
    body = np.frombuffer(gz, np.uint8)[deflate_off:-8].copy()
    for bit, out, win in [pts[k] for k in sorted(set(rng.integers(0, len(pts), 4).tolist()))]:
        w = np.frombuffer(win, np.uint8)
        have = min(out, 32768)
        assert np.array_equal(w[32768 - have:], raw[out - have:out])
        # the oracle resumes there with that window and reproduces the bytes
        want = raw[out:out + 4096]
        res, got, _ = oracle.inflate(body, bit, 2**64 - 1, window=w.tobytes(), out_cap=1 << 20)
        assert got.size >= want.size and np.array_equal(got[:want.size], want)
    for off, ln in [(0, 100), (raw.size - 5000, 10000), (int(raw.size * 0.37), 3 << 20), (outs[3], 1),
                    (outs[3] - 1, 2), (int(rng.integers(0, raw.size)), int(rng.integers(1, 1 << 22)))]:
        got = gpu.seek_decompress(gz, ix, off, ln)
        assert got == raw[off:off + ln].tobytes(), (off, ln)
    assert gpu.seek_decompress(gz, ix, 0) == raw.tobytes()


def test_index_refusals(gpu):
    raw = W.generate("text", 4 << 20, 3)
    gz = W.gzip_compress(raw, 6).tobytes()
    other = W.gzip_compress(W.generate("text", 4 << 20, 4), 6).tobytes()
    ix = gpu.build_index(gz, 256 << 10)
    with pytest.raises(gpu.GzippyError):
        gpu.build_index(gz + gz, 1 << 20)  # multi-member
    with pytest.raises(gpu.GzippyError):
        gpu.build_index(b"not gzip at all", 1 << 20)
    with pytest.raises(gpu.GzippyError):
        gpu.seek_decompress(gz, ix, raw.size + 1, 10)
    with pytest.raises(gpu.GzippyError):
        gpu.seek_decompress(other, ix, 1 << 20, 100)
    with pytest.raises(gpu.GzippyError):
        gpu.seek_decompress(gz, b"GZIDX\x02" + ix[6:], 0, 10)
