"""Speculation hazards on the GPU path, restated from the reference's tests:

* forced misses: every n-th planned start one bit late (a test hook in the
  tuning), so the re-decode path runs, also for runs of consecutive misses
  far longer than any fixed round cap (chunk_fetcher.rs:2506-2729);
* BTYPE=01-heavy members (src/tests/routing.rs:1599-1672);
* phantom end-of-stream blocks (src/tests/phantom_eos_probe.rs:44-146);
* gzip of deflate data: real block headers inside stored payloads;
* empty first and interior members (src/tests/multi_member_chunked.rs:135-168);
* back-references at distance 32767 and 32768 across piece seams
  (src/tests/seam_crossing.rs:442-470);
* header fields FEXTRA / FNAME / FCOMMENT / FHCRC (gzip_format.rs:47);
* a seeded sweep of truncations and bit flips: every call returns an error
  or the exact bytes (the reference's fuzz/parallel_sm_roundtrip.rs contract).
"""
import struct
import zlib

import numpy as np
import pytest

from gzippy_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture
def tuned(gpu):
    """Set a tuning for the test; restore the context's on exit."""
    ctx = gpu.context()
    old = ctx.tuning()

    def set_(**kw):
        ctx.set_tuning(**kw)

    yield set_
    ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes, 0)


def gz_fixed(raw, level=6):
    """One member of fixed-Huffman blocks only (zlib Z_FIXED)."""
    co = zlib.compressobj(level, zlib.DEFLATED, 31, 9, zlib.Z_FIXED)
    return co.compress(bytes(raw)) + co.flush()


@pytest.mark.parametrize("every", [1, 2, 3, 7])
@pytest.mark.parametrize("kind,level", [("text", 6), ("u32", 1), ("runs", 9)])
def test_forced_misses(gpu, tuned, every, kind, level):
    """Every `every`-th planned start is one bit late: its predecessor's end
    misses it and the piece is decoded again from the true end.  At every=1
    all of the member's hundreds of starts miss in a row."""
    raw = W.generate(kind, (28 if every == 1 else 12) << 20, seed=101 + every)
    gz = W.gzip_compress(raw, level)
    tuned(chunk_bytes=32 << 10, pieces_per_chunk=8, miss_every=every)
    st = {}
    out = np.frombuffer(gpu.decompress(gz, stats=st), np.uint8)
    say("forced misses", kind, level, every, {k: st[k] for k in ("pieces", "redecoded", "inflate_launches")})
    assert np.array_equal(out, raw)
    assert st["redecoded"] > 0
    if every == 1:
        assert st["pieces"] > 64  # a run of misses far past the old 64-launch cap
    assert st["inflate_launches"] <= 6


@pytest.mark.parametrize("kind", ["text", "u32", "runs"])
def test_fixed_huffman_heavy(gpu, tuned, kind):
    """Fixed-Huffman blocks only: the block finder reports no dynamic
    headers, so the member is planned without hits; with forced misses the
    shifted starts also meet phantom BFINAL=1 fixed blocks."""
    raw = W.generate(kind, 6 << 20, seed=55)
    gz = gz_fixed(raw)
    assert gpu.decompress(gz) == raw.tobytes()
    tuned(chunk_bytes=32 << 10, pieces_per_chunk=8, miss_every=2)
    st = {}
    assert gpu.decompress(gz, stats=st) == raw.tobytes()
    say("fixed", kind, {k: st[k] for k in ("pieces", "redecoded", "inflate_launches")})


def test_phantom_eos_vectors(gpu):
    """The reference's synthetic phantom: an empty BFINAL=1 fixed block, a
    fake footer, then bytes that are not a gzip header.  As raw deflate the
    piece ends at bit 10 with the final block; inside a member's stored
    payload the same bytes are data, not a stream end."""
    import torch
    fixed_empty_bfinal = bytes([0x03, 0x00])
    footer = bytes([0xAA, 0xBB, 0xCC, 0xDD, 0x01, 0, 0, 0])
    buf = np.frombuffer(fixed_empty_bfinal + footer + bytes(100), np.uint8)
    m = torch.zeros(64, dtype=torch.int16, device="cuda")
    r = gpu.inflate_piece(torch.from_numpy(buf.copy()).cuda(), 0, 2**64 - 1, m)
    assert r.status == 0 and r.bfinal_hit == 1 and r.end_bit == 10 and r.out_len == 0
    payload = (fixed_empty_bfinal + footer + bytes(100)) * 20000 + bytes(W.generate("random", 1 << 20, 3))
    for level in (0, 1, 6):
        c = zlib.compressobj(level, zlib.DEFLATED, 31)
        gz = c.compress(payload) + c.flush()
        assert gpu.decompress(gz) == payload, level


def test_gzip_of_deflate_data(gpu, oracle):
    """A member whose payload is itself deflate data (a .gz inside a .gz):
    stored blocks whose payloads hold real dynamic block headers.  Starts
    planned there would all miss; the planner leaves them out."""
    inner_raw = W.generate("u32", 96 << 20, seed=8)
    inner = W.gzip_compress(inner_raw, 1)
    for level in (1, 6):
        gz = W.gzip_compress(inner, level)
        say("gzip of gzip", level, inner.size, "->", gz.size)
        st_o, want, _ = oracle.gunzip(gz)
        assert st_o == 0 and np.array_equal(want, inner)
        st = {}
        out = gpu.decompress(gz, stats=st)
        say("stats", {k: st[k] for k in ("pieces", "redecoded", "inflate_launches", "candidates")})
        assert out == inner.tobytes()
        assert st["inflate_launches"] <= 3


@pytest.mark.parametrize("layout", ["empty_first", "empty_interior", "empty_last", "all_empty"])
def test_empty_members(gpu, layout):
    a = W.generate("text", 512 << 10, 400)
    b = W.generate("u32", 512 << 10, 401)
    e = np.zeros(0, np.uint8)
    parts = {"empty_first": [e, a], "empty_interior": [a, e, b], "empty_last": [a, b, e],
             "all_empty": [e, e, e]}[layout]
    gz = b"".join(W.gzip_compress(p, 6).tobytes() for p in parts)
    want = b"".join(p.tobytes() for p in parts)
    assert gpu.decompress(gz) == want
    import io
    sink = io.BytesIO()
    assert gpu.decompress_to_writer(gz, sink) == len(want) and sink.getvalue() == want


def periodic(n, period, seed=0xA5A5):
    rng = np.random.default_rng(seed ^ period)
    base = rng.integers(0, 256, period, dtype=np.uint8)
    return np.resize(base, n)


class BitWriter:
    """RFC 1951 bit packing: values LSB first, Huffman codes MSB first."""

    def __init__(self):
        self.acc, self.n, self.out = 0, 0, bytearray()

    def bits(self, v, n):
        self.acc |= v << self.n
        self.n += n
        while self.n >= 8:
            self.out.append(self.acc & 255)
            self.acc >>= 8
            self.n -= 8

    def code(self, c, n):
        self.bits(int(format(c, "0%db" % n)[::-1], 2), n)

    def align(self):
        if self.n:
            self.bits(0, 8 - self.n)

    def lit(self, v):  # fixed literal/length code
        if v < 144:
            self.code(0x30 + v, 8)
        else:
            self.code(0x190 + v - 144, 9)

    def copy258(self, dist):  # length 258 (symbol 285), distance code 29
        self.code(0xC0 + 5, 8)
        assert 24577 <= dist <= 32768
        self.code(29, 5)
        self.bits(dist - 24577, 13)

    def stored(self, data, final=0):
        self.bits(final, 1)
        self.bits(0, 2)
        self.align()
        self.bits(len(data), 16)
        self.bits(len(data) ^ 0xFFFF, 16)
        for b in data:
            self.bits(b, 8)


def window_distance_stream(segments=48, seed=7):
    """Fixed-Huffman blocks of 258-byte copies at distances 32768 and 32767
    (which no zlib or libdeflate level emits) and literals, each followed by
    two stored blocks (a chained stored hit, where a piece can start): the
    copies of every block after a seam reach the window's oldest bytes."""
    rng = np.random.default_rng(seed)
    w = BitWriter()
    for s in range(segments):
        w.bits(0, 1)
        w.bits(1, 2)
        if s == 0:
            for v in rng.integers(0, 256, 32768):
                w.lit(int(v))
        for k in range(300):
            w.copy258(32768 if (k + s) % 3 else 32767)
            if k % 7 == 0:
                w.lit(int(rng.integers(0, 256)))
        w.code(0, 7)  # end of block
        w.stored(bytes(rng.integers(0, 256, 100, dtype=np.uint8)))
        w.stored(bytes(rng.integers(0, 256, 60, dtype=np.uint8)))
    w.bits(1, 1)
    w.bits(1, 2)
    w.code(0, 7)
    w.align()
    return bytes(w.out)


@pytest.mark.parametrize("form", ["libdeflate-32767", "hand-built-32768"])
def test_seam_window_distances(gpu, tuned, form):
    """Copies at distance 32767 / 32768 (the window's extremities) across
    piece seams.  libdeflate emits distances up to 32767 (period-32767 data);
    32768 needs a hand-built stream."""
    if form == "libdeflate-32767":
        raw = periodic(8 << 20, 32767)
        gz = W.gzip_compress(raw, 6, encoder="libdeflate")
        assert gz.size < raw.size // 20  # the matches were found
        gz = gz.tobytes()
        raw = raw.tobytes()
    else:
        d = window_distance_stream()
        raw = zlib.decompress(d, -15)
        gz = b"\x1f\x8b\x08\x00\x00\x00\x00\x00\x00\xff" + d + struct.pack("<II", zlib.crc32(raw), len(raw))
    tuned(chunk_bytes=64 << 10, pieces_per_chunk=8)
    st = {}
    assert gpu.decompress(gz, stats=st) == raw
    say("seam", form, len(raw), {k: st[k] for k in ("pieces", "redecoded")})
    assert st["pieces"] > 4


def gz_with_fields(body_raw, level=6, fextra=None, fname=None, fcomment=None, fhcrc=False):
    """A member with the optional RFC 1952 header fields."""
    flg = (4 if fextra is not None else 0) | (8 if fname is not None else 0) | \
        (16 if fcomment is not None else 0) | (2 if fhcrc else 0)
    h = bytearray(b"\x1f\x8b\x08" + bytes([flg]) + b"\x00\x00\x00\x00\x00\xff")
    if fextra is not None:
        h += struct.pack("<H", len(fextra)) + fextra
    if fname is not None:
        h += fname + b"\x00"
    if fcomment is not None:
        h += fcomment + b"\x00"
    if fhcrc:
        h += struct.pack("<H", zlib.crc32(bytes(h)) & 0xFFFF)
    co = zlib.compressobj(level, zlib.DEFLATED, -15)
    body = co.compress(bytes(body_raw)) + co.flush()
    return bytes(h) + body + struct.pack("<II", zlib.crc32(bytes(body_raw)), len(body_raw) & 0xFFFFFFFF)


def test_header_fields(gpu):
    import torch
    raw = W.generate("text", 3 << 20, 77)
    cases = {
        "fname": dict(fname=b"data.txt"),
        "fcomment": dict(fcomment=b"a comment " * 10),
        "fhcrc": dict(fhcrc=True),
        "all": dict(fextra=b"ab\x04\x00wxyz", fname=b"n" * 300, fcomment=b"c" * 1000, fhcrc=True),
        # a FEXTRA near its 65535-byte limit followed by FNAME: the header
        # runs past the first 64 KiB the device path reads
        "long": dict(fextra=b"XY" + struct.pack("<H", 65531) + bytes(65531), fname=b"x" * 5000),
    }
    for name, kw in cases.items():
        gz = gz_with_fields(raw, **kw)
        say("header", name, len(gz))
        assert gpu.decompress(gz) == raw.tobytes(), name
        out = torch.zeros(raw.size, dtype=torch.uint8, device="cuda")
        n, crc, _ = gpu.decompress_device(torch.from_numpy(np.frombuffer(gz, np.uint8).copy()).cuda(), out)
        assert n == raw.size and crc == zlib.crc32(raw.tobytes()), name
        assert np.array_equal(out.cpu().numpy(), raw), name
    # a reserved flag bit or a bad method is an invalid header
    gz = bytearray(gz_with_fields(raw[:1000]))
    gz[2] = 7
    with pytest.raises(gpu.GzippyError) as e:
        gpu.decompress(bytes(gz))
    assert e.value.kind == "InvalidHeader"


def test_corruption_sweep(gpu, tuned):
    """Seeded truncations and bit flips of a 1 MiB member (split into pieces):
    every call either raises GzippyError or returns the original bytes."""
    raw = W.generate("text", 1 << 20, 99)
    gz = W.gzip_compress(raw, 6)
    tuned(chunk_bytes=16 << 10, pieces_per_chunk=4)
    rng = np.random.default_rng(12345)
    outcomes = {"error": 0, "exact": 0}
    for case in range(120):
        bad = gz.copy()
        if case % 3 == 0:
            bad = bad[: int(rng.integers(1, gz.size))]
            what = f"truncate {bad.size}"
        else:
            nflip = 1 + case % 4
            pos = rng.integers(0, gz.size, nflip)
            bits = rng.integers(0, 8, nflip)
            for p, b in zip(pos, bits):
                bad[p] ^= np.uint8(1 << int(b))
            what = f"flip {list(zip(pos.tolist(), bits.tolist()))}"
        say("corrupt case", case, what)
        try:
            out = gpu.decompress(bad)
        except gpu.GzippyError:
            outcomes["error"] += 1
            continue
        # a flip in the header's unused fields (MTIME, XFL, OS) or in bits a
        # block never reads leaves the stream valid
        assert out == raw.tobytes() or out == b"", what
        outcomes["exact"] += 1
    say("outcomes", outcomes)
    assert outcomes["error"] > 60
