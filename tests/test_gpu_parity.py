"""GPU parity: every kernel against the oracle on the same seeded inputs,
through the C ABI.  Integer/byte work: bit-exact everywhere."""
import hashlib
import os
import zlib

import numpy as np
import pytest

from gzippy_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "silesia_blocks.npz")


def say(*a):
    print("[gpu-test]", *a, flush=True)


def dev(arr):
    import torch
    return torch.from_numpy(np.ascontiguousarray(arr)).cuda()


def deflate_raw(data, level=6):
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 8)
    return np.frombuffer(c.compress(bytes(data)) + c.flush(), np.uint8)


# ---------------------------------------------------------------- CRC32

@pytest.mark.parametrize("n", [0, 1, 255, 256, 257, 4095, 65535, 65536, 65537, (1 << 20) + 13, (16 << 20) + 5])
def test_crc32_device(gpu, n):
    say("crc32 n=", n)
    rng = np.random.default_rng(n)
    d = rng.integers(0, 256, max(n, 1), dtype=np.uint8)[:n]
    t = dev(d if n else np.zeros(1, np.uint8))
    assert gpu.crc32(t, n) == zlib.crc32(d.tobytes())


def test_crc32_unaligned(gpu):
    """Unaligned starts and ends: a segment that starts or ends inside a
    16-byte chunk is staged with masked 16-byte loads (down to one chunk
    holding the whole segment) and the trailing pad's shift divided out."""
    rng = np.random.default_rng(7)
    d = rng.integers(0, 256, 300_001, dtype=np.uint8)
    t = dev(d)
    for off in (1, 2, 3, 5, 8, 15):
        sub = t[off:]
        assert gpu.crc32(sub, d.size - off) == zlib.crc32(d[off:].tobytes())
        for n in (1, 7, 14, 16, 17, 4095, 65535 - off, 65536, 65537 + off):
            assert gpu.crc32(sub, n) == zlib.crc32(d[off:off + n].tobytes()), (off, n)


# ---------------------------------------------------------------- block finder

FINDER_CASES = [("text", 6, "libdeflate", 2 << 20), ("text", 1, "zlib", 1 << 20), ("u32", 9, "zlib", 1 << 20),
                ("random", 6, "libdeflate", 1 << 20), ("runs", 6, "zlib", 1 << 20), ("random", 1, "zlib", 300_000)]


@pytest.mark.parametrize("kind,level,enc,size", FINDER_CASES)
def test_find_blocks_equals_oracle(gpu, oracle, kind, level, enc, size):
    say("find_blocks", kind, level, enc, size)
    raw = W.generate(kind, size, seed=101)
    gz = W.gzip_compress(raw, level, encoder=enc)
    want = oracle.find_blocks(gz, 0, gz.size * 8)
    got = gpu.find_blocks(dev(gz), 0, gz.size * 8)
    assert got == want
    # a sub-range with unaligned ends
    lo, hi = 12345, gz.size * 8 - 777
    assert gpu.find_blocks(dev(gz), lo, hi) == oracle.find_blocks(gz, lo, hi)


def test_find_blocks_recall_true_boundaries(gpu, oracle):
    raw = W.generate("text", 4 << 20, seed=5)
    gz = W.gzip_compress(raw, 6)
    st, hl = oracle.parse_header(gz)
    res, _, blocks = oracle.inflate(gz[hl:-8], 0, blocks_cap=1 << 16)
    hits = {h[0] for h in gpu.find_blocks(dev(gz))}
    true_dyn = [hl * 8 + b[0] for b in blocks[1:] if b[2] == 2 and not b[3]]
    assert len(true_dyn) > 8
    assert all(b in hits for b in true_dyn)


# ---------------------------------------------------------------- one piece

def test_inflate_piece_silesia_vectors(gpu, oracle):
    import torch
    z = np.load(GOLD)
    for i, (idx, btype, bitoff, bitlen, declen, crc) in enumerate(z["meta"]):
        comp = np.concatenate([z[f"comp{i}"], np.zeros(8, np.uint8)])
        win = z[f"win{i}"]
        out = torch.zeros(int(declen) + 4096, dtype=torch.uint8, device="cuda")
        r = gpu.inflate_piece(dev(comp), int(bitoff), int(bitoff) + 1, out, window=win.tobytes())
        assert r.status == 0 and r.out_len == declen and r.end_bit - bitoff == bitlen, i
        got = out[: int(declen)].cpu().numpy()
        assert hashlib.sha256(got.tobytes()).digest() == z[f"sha{i}"].tobytes(), i
        # window-absent: u16 markers, resolved on the CPU by the oracle
        m = torch.zeros(int(declen) + 4096, dtype=torch.int16, device="cuda")
        r = gpu.inflate_piece(dev(comp), int(bitoff), int(bitoff) + 1, m)
        assert r.status == 0 and r.out_len == declen, i
        marks = m[: int(declen)].cpu().numpy().view(np.uint16)
        st, resolved = oracle.resolve_markers(marks, win)
        assert st == 0 and np.array_equal(resolved, got), i


@pytest.mark.parametrize("kind,level", [("text", 6), ("u32", 1), ("runs", 9), ("random", 6), ("text", 9)])
def test_inflate_piece_markers_equal_oracle(gpu, oracle, kind, level):
    import torch
    say("piece markers", kind, level)
    raw = W.generate(kind, 3 << 20, seed=33)
    d = deflate_raw(raw, level)
    res, _, blocks = oracle.inflate(d, 0, blocks_cap=1 << 16)
    td = dev(d)
    for bi in range(1, min(len(blocks), 5)):
        bit, off = blocks[bi][0], blocks[bi][1]
        stop = blocks[bi + 2][0] if bi + 2 < len(blocks) else 2**64 - 1
        ores, omarks, _ = oracle.inflate(d, bit, stop, marker_mode=True)
        m = torch.zeros(int(ores.out_len) + 8192, dtype=torch.int16, device="cuda")
        r = gpu.inflate_piece(td, bit, stop, m)
        assert r.status == 0 and r.out_len == ores.out_len and r.end_bit == ores.end_bit
        got = m[: r.out_len].cpu().numpy().view(np.uint16)
        assert np.array_equal(got, omarks), (kind, level, bi)


def long_round_stream(run_len, reps=120, seed=5):
    """Runs of one byte long enough that a 64-bit lookahead window holds a
    ring's worth of output (many 258-byte copies with 2-3 bit codes), each
    followed by literals that later copies reference at short distance: a
    round whose copy destinations alias its own later literals in the ring."""
    rng = np.random.default_rng(seed)
    parts = []
    for i in range(reps):
        a, b = rng.integers(65, 91, 2)
        tail = bytes([b, (b + 1) % 256, (b + 2) % 256]) * 4
        parts.append(bytes([a]) * run_len + tail + bytes(rng.integers(0, 256, 3, dtype=np.uint8)))
    return np.frombuffer(b"".join(parts), np.uint8).copy()


@pytest.mark.parametrize("run_len", [4500, 7740, 9100, 16500, 33000])
def test_round_longer_than_ring(gpu, oracle, run_len):
    """Both kernels on rounds that produce more entries than their LDS ring:
    marker mode (u16 ring) and window mode (u8 ring) from the stream start,
    and the whole member through the pipeline."""
    import torch
    raw = long_round_stream(run_len)
    say("long rounds", run_len, raw.size)
    for level in (6, 9):
        d = deflate_raw(raw, level)
        td = dev(d)
        m = torch.zeros(raw.size + 8192, dtype=torch.int16, device="cuda")
        r = gpu.inflate_piece(td, 0, 2**64 - 1, m)
        assert r.status == 0 and r.out_len == raw.size
        got = m[: raw.size].cpu().numpy().view(np.uint16)
        assert np.array_equal(got, raw.astype(np.uint16)), (run_len, level, "marker")
        o = torch.zeros(raw.size + 64, dtype=torch.uint8, device="cuda")
        r = gpu.inflate_piece(td, 0, 2**64 - 1, o, window=bytes(32768))
        assert r.status == 0 and r.out_len == raw.size
        assert np.array_equal(o[: raw.size].cpu().numpy(), raw), (run_len, level, "window")
        gz = W.gzip_compress(raw, level)
        assert gpu.decompress(gz) == raw.tobytes(), (run_len, level, "member")


def test_runs_level9_split_members(gpu):
    """Long-run data at level 9 split into speculative pieces (the marker
    kernel): copies far back and rounds of several thousand entries."""
    ctx = gpu.context()
    old = ctx.tuning()
    ctx.set_tuning(chunk_bytes=64 << 10, pieces_per_chunk=4)
    try:
        for seed in (9113, 77, 4242):
            raw = W.generate("runs", 8 << 20, seed)
            gz = W.gzip_compress(raw, 9)
            st = {}
            assert gpu.decompress(gz, stats=st) == raw.tobytes(), seed
            say("runs L9", seed, gz.size, "pieces", st["pieces"])
    finally:
        ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes)
