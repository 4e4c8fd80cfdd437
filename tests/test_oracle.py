"""Pins the CPU oracle (oracle/gzoracle.c) before it is trusted as the checker.

* decode: byte equality with the generator's own bytes and with two
  independent decoders (system zlib, libdeflate), CRC32/ISIZE trailers;
* golden vectors: the reference's 30 frozen GZBLK01 blocks
  (corpus/silesia_blocks, fixture tests/golden/silesia_blocks.npz);
* CRC32: known answer, zlib crc32 / crc32_combine;
* markers: window-absent decode + MapMarkers resolution == straight decode
  (the seam_crossing.rs pattern, src/tests/seam_crossing.rs:64-310);
* block finder: 100 % recall of the true non-final dynamic and stored
  boundaries; every dynamic hit re-parses as a valid header through the
  inflater's own (independent) header reader; constructed headers probe each
  rule of blockfinder_validation.rs.
"""
import ctypes
import hashlib
import os
import zlib

import numpy as np
import pytest

from gzippy_amd import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden", "silesia_blocks.npz")
_z = ctypes.CDLL("libz.so.1")
_z.crc32_combine.restype = ctypes.c_ulong
_z.crc32_combine.argtypes = [ctypes.c_ulong, ctypes.c_ulong, ctypes.c_long]


def deflate_raw(data, level=6, strategy=zlib.Z_DEFAULT_STRATEGY):
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 8, strategy)
    return c.compress(bytes(data)) + c.flush()


@pytest.mark.parametrize("kind", ["text", "random", "runs", "u32"])
@pytest.mark.parametrize("level", [1, 6, 9])
def test_gunzip_matches_generator(oracle, kind, level):
    raw = W.generate(kind, 300_000 + 17 * level, seed=11 + level)
    for enc in ("libdeflate", "zlib"):
        gz = W.gzip_compress(raw, level, encoder=enc)
        st, out, members = oracle.gunzip(gz)
        assert st == oracle.OK and members == 1
        assert out.tobytes() == raw.tobytes()
        assert zlib.decompress(bytes(gz), 31) == raw.tobytes()


def test_gunzip_text_member_shape(oracle):
    raw, gz = W.text_member(16 << 20, seed=0x675A0000)
    st, out, _ = oracle.gunzip(gz)
    assert st == oracle.OK and np.array_equal(out, raw)
    crc, isize = W.trailer(gz)
    assert crc == zlib.crc32(raw.tobytes()) and isize == raw.size


def test_gunzip_multi_member_and_garbage(oracle):
    a, b = W.generate("text", 50_000, 1), W.generate("u32", 70_001, 2)
    gz = np.concatenate([W.gzip_compress(a, 6), W.gzip_compress(b, 1)])
    st, out, m = oracle.gunzip(gz)
    assert st == 0 and m == 2 and out.tobytes() == a.tobytes() + b.tobytes()
    st, out, m = oracle.gunzip(np.concatenate([gz, np.frombuffer(b"trailing junk", np.uint8)]))
    assert st == 0 and m == 2 and out.tobytes() == a.tobytes() + b.tobytes()


def test_gunzip_rejects_corruption(oracle):
    raw = W.generate("text", 200_000, 5)
    gz = W.gzip_compress(raw, 6).copy()
    bad = gz.copy()
    bad[-6] ^= 1  # CRC byte
    assert oracle.gunzip(bad)[0] == oracle.CRC_MISMATCH
    bad = gz.copy()
    bad[-2] ^= 1  # ISIZE byte
    assert oracle.gunzip(bad)[0] == oracle.SIZE_MISMATCH
    assert oracle.gunzip(gz[: gz.size // 2])[0] == oracle.DECODE_FAILED
    assert oracle.gunzip(b"not gzip at all")[0] == oracle.OK  # non-gzip: Ok(0)
    assert oracle.gunzip(b"\x1f\x8b\x07")[0] == oracle.INVALID_HEADER


def test_header_fields(oracle):
    import gzip
    raw = b"hello header fields" * 100
    # FNAME + FEXTRA + FCOMMENT + FHCRC built by hand around a raw deflate body
    body = deflate_raw(raw)
    hdr = bytearray(b"\x1f\x8b\x08\x1e" + b"\0\0\0\0\0\x03")
    hdr += (5).to_bytes(2, "little") + b"EXTRA" + b"name.txt\0" + b"a comment\0"
    hdr += (zlib.crc32(bytes(hdr)) & 0xFFFF).to_bytes(2, "little")
    gz = bytes(hdr) + body + zlib.crc32(raw).to_bytes(4, "little") + len(raw).to_bytes(4, "little")
    st, hl = oracle.parse_header(gz)
    assert st == 0 and hl == len(hdr)
    st, out, _ = oracle.gunzip(gz)
    assert st == 0 and out.tobytes() == raw
    assert gzip.decompress(gz) == raw


def test_sized_member_streams(oracle):
    """gzippy's "GZ" members and BGZF members (with its empty end-of-file
    member) are ordinary multi-member streams to the oracle."""
    import gzip
    raw = W.generate_mixed(2 << 20, 3, seg=256 << 10)
    for form, block in (("gz", 300_000), ("bgzf", 65280)):
        gz = W.sized_member_stream(raw, block, 6, form)
        assert gzip.decompress(gz) == raw.tobytes()
        st, out, _ = oracle.gunzip(gz)
        assert st == oracle.OK and out.tobytes() == raw.tobytes(), form


# ---------------------------------------------------------------- golden vectors

def test_silesia_golden_vectors(oracle):
    z = np.load(GOLD)
    meta = z["meta"]
    assert len(meta) == 30
    for i, (idx, btype, bitoff, bitlen, declen, crc) in enumerate(meta):
        # the corpus keeps only the bytes covering the block's bit range; zero
        # padding makes the bits after it read as a non-final stored header
        comp = np.concatenate([z[f"comp{i}"], np.zeros(8, np.uint8)])
        win = z[f"win{i}"]
        # one block: any following non-final header stops the decode
        res, out, blocks = oracle.inflate(comp, int(bitoff), int(bitoff) + 1, window=win,
                                          out_cap=int(declen) + 1024, blocks_cap=2)
        assert res.status == 0 and res.out_len == declen
        assert res.end_bit - bitoff == bitlen
        assert hashlib.sha256(out.tobytes()).digest() == z[f"sha{i}"].tobytes()
        assert oracle.crc32(out) == crc == zlib.crc32(out.tobytes())
        assert blocks[0][2] == btype == 2
        # marker mode + MapMarkers against the true window
        res2, marks, _ = oracle.inflate(comp, int(bitoff), int(bitoff) + 1, marker_mode=True,
                                        out_cap=int(declen) + 1024)
        st, resolved = oracle.resolve_markers(marks, win)
        assert st == 0 and np.array_equal(resolved, out)
        # the block start is a block-finder candidate
        assert oracle.is_dynamic_candidate(comp, int(bitoff))


# ---------------------------------------------------------------- CRC32

def test_crc32_known_answers(oracle):
    assert oracle.crc32(b"123456789") == 0xCBF43926
    assert oracle.crc32(b"") == 0
    rng = np.random.default_rng(3)
    for n in [1, 7, 8, 9, 63, 64, 65, 1000, 70_001]:
        d = rng.integers(0, 256, n, dtype=np.uint8)
        for off in range(4):
            assert oracle.crc32(d[off:]) == zlib.crc32(d[off:].tobytes())


def test_crc32_combine_matches_zlib(oracle):
    rng = np.random.default_rng(4)
    for la, lb in [(0, 5), (5, 0), (1, 1), (100, 1000), (65536, 3), (12345, 98765)]:
        a = rng.integers(0, 256, la, dtype=np.uint8).tobytes()
        b = rng.integers(0, 256, lb, dtype=np.uint8).tobytes()
        ca, cb = zlib.crc32(a), zlib.crc32(b)
        want = zlib.crc32(a + b)
        assert oracle.crc32_combine(ca, cb, lb) == want == _z.crc32_combine(ca, cb, lb)


# ---------------------------------------------------------------- markers

def _true_blocks(oracle, raw_deflate):
    res, out, blocks = oracle.inflate(raw_deflate, 0, blocks_cap=1 << 16)
    assert res.status == 0 and res.bfinal_hit
    return out, blocks


@pytest.mark.parametrize("kind,level", [("text", 6), ("u32", 1), ("runs", 9), ("random", 6)])
def test_marker_decode_resolves_to_straight_decode(oracle, kind, level):
    raw = W.generate(kind, 3 << 20, seed=21)
    d = np.frombuffer(deflate_raw(raw, level), np.uint8)
    out, blocks = _true_blocks(oracle, d)
    assert np.array_equal(out, raw)
    assert len(blocks) >= 3
    for bi in range(1, min(len(blocks), 6)):
        bit, off, btype, bfinal = blocks[bi]
        stop = blocks[bi + 2][0] if bi + 2 < len(blocks) else 2**64 - 1
        res, marks, _ = oracle.inflate(d, bit, stop, marker_mode=True)
        assert res.status == 0
        window = np.zeros(32768, np.uint8)
        lo = max(0, off - 32768)
        window[32768 - (off - lo):] = raw[lo:off]
        st, resolved = oracle.resolve_markers(marks, window)
        assert st == 0
        assert np.array_equal(resolved, raw[off: off + res.out_len])
        if bi + 2 < len(blocks) and not blocks[bi + 2][3]:
            assert res.end_bit == blocks[bi + 2][0]  # stops exactly at the stop block
        else:
            assert res.bfinal_hit  # a final block is decoded, never stopped at


def test_marker_values_distance_32768(oracle):
    # a back-reference of exactly 32768 from the piece start hits window[0]
    raw = bytearray(os.urandom(32768)) + bytearray(os.urandom(40000))
    raw[32768 + 1000: 32768 + 1300] = raw[1000:1300]
    d = np.frombuffer(deflate_raw(bytes(raw), 9), np.uint8)
    out, blocks = _true_blocks(oracle, d)
    assert out.tobytes() == bytes(raw)


# ---------------------------------------------------------------- block finder

def _check_recall(oracle, gzbytes, ds_bits, blocks):
    hits = oracle.find_blocks(gzbytes, ds_bits + 1, len(gzbytes) * 8)
    hit_bits = {h[0] for h in hits}
    hit_ranges = [(h[0], h[1]) for h in hits]
    for bit, off, btype, bfinal in blocks[1:]:
        abs_bit = ds_bits + bit
        if bfinal:
            continue
        if btype == 2:
            assert abs_bit in hit_bits, f"dynamic boundary {abs_bit} missed"
        elif btype == 0:
            assert any(lo <= abs_bit <= hi for lo, hi in hit_ranges), f"stored boundary {abs_bit} missed"
    return hits


@pytest.mark.parametrize("kind,level,enc", [("text", 6, "libdeflate"), ("text", 1, "zlib"), ("u32", 9, "zlib"),
                                            ("random", 6, "libdeflate"), ("runs", 6, "zlib")])
def test_block_finder_recall(oracle, kind, level, enc):
    raw = W.generate(kind, 2 << 20, seed=31)
    gz = W.gzip_compress(raw, level, encoder=enc)
    st, hl = oracle.parse_header(gz)
    body = gz[hl:-8]
    _, blocks = _true_blocks(oracle, body)
    hits = _check_recall(oracle, gz, hl * 8, blocks)
    # every dynamic hit re-parses through the inflater's header reader
    for bo, sb, k in hits[:400]:
        if k == 2:
            res, _, blk = oracle.inflate(gz, bo, bo + 1, marker_mode=True, out_cap=1 << 22, blocks_cap=1)
            assert res.n_blocks >= 1 and blk[0][0] == bo and blk[0][2] == 2


class BitWriter:
    def __init__(self):
        self.bits = []

    def put(self, v, n):
        for i in range(n):
            self.bits.append((v >> i) & 1)

    def put_code(self, code, n):  # Huffman code, MSB first
        for i in reversed(range(n)):
            self.bits.append((code >> i) & 1)

    def bytes(self, pad=256):
        b = self.bits + [0] * ((-len(self.bits)) % 8) + [0] * pad
        return np.packbits(np.array(b, np.uint8), bitorder="little")


PRE_ORDER = [16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15]


def _header(bfinal=0, btype=2, hlit=0, hdist=0, pre_lens=None, body=(), hclen_count=18):
    """Dynamic block header bits.  `pre_lens` maps precode symbol -> length
    (RFC 1951 3.2.7); `body` lists (precode symbol, extra value, extra bits)."""
    w = BitWriter()
    w.put(bfinal, 1)
    w.put(btype, 2)
    w.put(hlit, 5)
    w.put(hdist, 5)
    w.put(hclen_count - 4, 4)
    for i in range(hclen_count):
        w.put(pre_lens.get(PRE_ORDER[i], 0), 3)
    codes, code = {}, 0  # canonical code assignment (RFC 1951 3.2.2)
    for L in range(1, 8):
        for s in range(19):
            if pre_lens.get(s, 0) == L:
                codes[s] = (code, L)
                code += 1
        code <<= 1
    for sym, extra, nb in body:
        c, L = codes[sym]
        w.put_code(c, L)
        if nb:
            w.put(extra, nb)
    return w.bytes()


# precode: symbols 0, 1, 18 of length 2 and 16, 2 of length 3 (Kraft sum 1)
PRE = {0: 2, 1: 2, 18: 2, 16: 3, 2: 3}
# litlen (257 codes): sym 0 len 1, 255 zeros, EOB len 1; distance: one 1-bit code
VALID = [(1, 0, 0), (18, 127, 7), (18, 106, 7), (1, 0, 0), (1, 0, 0)]


def test_block_finder_constructed_headers(oracle):
    cand = lambda b: oracle.is_dynamic_candidate(b, 0)
    assert cand(_header(pre_lens=PRE, body=VALID))
    assert not cand(_header(bfinal=1, pre_lens=PRE, body=VALID))          # BFINAL must be 0
    assert not cand(_header(btype=1, pre_lens=PRE, body=VALID))           # BTYPE must be 10
    assert not cand(_header(hlit=30, pre_lens=PRE, body=VALID))           # HLIT <= 29
    assert not cand(_header(hdist=30, pre_lens=PRE, body=VALID))          # HDIST <= 29
    # precode Kraft sum 7/8: leaf count 112
    assert not cand(_header(pre_lens={0: 2, 1: 2, 18: 2, 2: 3}, body=VALID))
    # EOB length 0 (litlen is then the single 1-bit code, still rejected)
    assert not cand(_header(pre_lens=PRE, body=[(1, 0, 0), (18, 127, 7), (18, 106, 7), (0, 0, 0), (1, 0, 0)]))
    # repeat-previous as the first code length
    assert not cand(_header(pre_lens=PRE, body=[(16, 0, 2)] + VALID))
    # repeat overflowing the code count
    assert not cand(_header(pre_lens=PRE, body=[(1, 0, 0), (18, 127, 7), (18, 127, 7)]))
    # litlen Kraft sum 1/2 from two length-2 codes: not the single-code case
    assert not cand(_header(pre_lens=PRE, body=[(2, 0, 0), (18, 127, 7), (18, 106, 7), (2, 0, 0), (1, 0, 0)]))
    # distance: a single code of length 2 (sum 1/4)
    assert not cand(_header(pre_lens=PRE, body=[(1, 0, 0), (18, 127, 7), (18, 106, 7), (1, 0, 0), (2, 0, 0)]))
    # distance: two 1-bit codes (complete) is valid
    assert cand(_header(hdist=1, pre_lens=PRE, body=VALID + [(1, 0, 0)]))
    # the header must start 17 bits before the end of the data
    b = _header(pre_lens=PRE, body=VALID)
    assert not oracle.is_dynamic_candidate(b[:2], 0)


def test_block_finder_stored(oracle):
    # stored block of 5 bytes preceded by a byte of zeros at a known offset
    data = bytearray(b"\xff" * 3) + bytearray(b"\x00") + (5).to_bytes(2, "little") + (0xFFFF ^ 5).to_bytes(2, "little") + b"hello" + b"\xaa" * 16
    hits = oracle.find_blocks(np.frombuffer(bytes(data), np.uint8), 0, len(data) * 8)
    stored = [h for h in hits if h[2] == 0]
    # LEN at byte 4: 8 zero bits before it (byte 3) -> earliest = 32 - 8 = 24, seek = 29
    assert (24, 29, 0) in stored


@pytest.mark.parametrize("kind", ["text", "random", "runs", "u32"])
@pytest.mark.parametrize("level", [1, 6, 9])
def test_parallel_restatement_matches_generator(oracle, kind, level):
    """The threaded speculative restatement (the bench's multi-core CPU
    baseline) returns the generator's bytes, with small partitions so that
    every member is split and some partitions start on false candidates."""
    raw = W.generate(kind, (2 << 20) + 4321, seed=40 + level)
    gz = W.gzip_compress(raw, level)
    for threads in (1, 3):
        st, out = oracle.gunzip_parallel(gz, threads=threads, chunk_bytes=32 << 10)
        assert st == 0 and np.array_equal(out, raw), (kind, level, threads)
    bad = gz.copy()
    bad[-7] ^= 1  # CRC
    st, _ = oracle.gunzip_parallel(bad, threads=2, chunk_bytes=32 << 10)
    assert st == oracle.CRC_MISMATCH
