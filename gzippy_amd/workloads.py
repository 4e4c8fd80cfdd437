"""Seeded synthetic gzip workloads.

Content comes from ``csrc/synth.c`` (four kinds: text, random, runs, u32
deltas); members are encoded with libdeflate (``libdeflate.so.0``) when it is
present, else zlib.  gzippy's own Rust encoder is not available here, so the
encoder used is reported with every workload; decoded bytes, CRC32 and ISIZE
do not depend on the encoder.
"""
import ctypes
import ctypes.util
import os
import struct
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import build as _build

KINDS = {"text": 0, "random": 1, "runs": 2, "u32": 3}
SEG = 64 << 20  # mixed-entropy segment size

_synth = None


def _synth_lib():
    global _synth
    if _synth is None:
        path = _build.synth_path()
        if not os.path.exists(path):
            _build.build_synth()
        L = ctypes.CDLL(path)
        L.gzs_generate.restype = ctypes.c_int
        L.gzs_generate.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64]
        _synth = L
    return _synth


def generate(kind, n, seed):
    """n bytes of `kind` ("text" | "random" | "runs" | "u32") from `seed`."""
    out = np.empty(int(n), dtype=np.uint8)
    if n:
        st = _synth_lib().gzs_generate(KINDS[kind], out.ctypes.data, int(n), int(seed))
        if st:
            raise ValueError(kind)
    return out


def generate_mixed(n, seed, seg=SEG, threads=8):
    """Cycle `seg`-byte segments of text / random / runs / u32 (one seed each)."""
    out = np.empty(int(n), dtype=np.uint8)
    order = ["text", "random", "runs", "u32"]
    jobs = []
    for i, off in enumerate(range(0, int(n), seg)):
        jobs.append((order[i % 4], off, min(seg, int(n) - off), seed * 1000003 + i))

    def one(job):
        kind, off, ln, s = job
        view = out[off:off + ln]
        _synth_lib().gzs_generate(KINDS[kind], view.ctypes.data, ln, s)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(one, jobs))
    return out


# ---------------------------------------------------------------- encoders

_ld = None


def _libdeflate():
    global _ld
    if _ld is None:
        try:
            L = ctypes.CDLL("libdeflate.so.0")
        except OSError:
            _ld = False
            return _ld
        L.libdeflate_alloc_compressor.restype = ctypes.c_void_p
        L.libdeflate_alloc_compressor.argtypes = [ctypes.c_int]
        L.libdeflate_free_compressor.argtypes = [ctypes.c_void_p]
        L.libdeflate_gzip_compress.restype = ctypes.c_size_t
        L.libdeflate_gzip_compress.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                               ctypes.c_void_p, ctypes.c_size_t]
        L.libdeflate_gzip_compress_bound.restype = ctypes.c_size_t
        L.libdeflate_gzip_compress_bound.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        _ld = L
    return _ld


def encoder_name():
    return "libdeflate" if _libdeflate() else "zlib-" + zlib.ZLIB_RUNTIME_VERSION


def gzip_compress(data, level=6, encoder=None):
    """One RFC 1952 member of `data` at `level` (libdeflate, else zlib)."""
    a = np.ascontiguousarray(np.frombuffer(data, dtype=np.uint8) if not isinstance(data, np.ndarray) else data)
    L = _libdeflate() if encoder in (None, "libdeflate") else False
    if L:
        c = L.libdeflate_alloc_compressor(int(level))
        try:
            bound = L.libdeflate_gzip_compress_bound(c, a.size)
            out = np.empty(bound, dtype=np.uint8)
            n = L.libdeflate_gzip_compress(c, a.ctypes.data, a.size, out.ctypes.data, bound)
            if n == 0:
                raise RuntimeError("libdeflate_gzip_compress failed")
            return out[:n].copy()
        finally:
            L.libdeflate_free_compressor(c)
    co = zlib.compressobj(int(level), zlib.DEFLATED, 31)
    body = co.compress(memoryview(a)) + co.flush()
    return np.frombuffer(body, dtype=np.uint8).copy()


def _crc_mulmod(a, b):
    """a * b mod P in the reflected CRC-32 polynomial basis."""
    r = 0
    m = 1 << 31
    while m:
        if a & m:
            r ^= b
        b = (b >> 1) ^ 0xEDB88320 if b & 1 else b >> 1
        m >>= 1
    return r


def crc32_combine(crc_a, crc_b, len_b):
    """CRC32 of A || B from crc(A), crc(B) and |B|: crc_a * x^(8 |B|) mod P,
    xor crc_b (the standard combine; x^k by square-and-multiply)."""
    p = 1 << 31  # x^0
    sq = 1 << 30  # x^1
    k = 8 * len_b
    while k:
        if k & 1:
            p = _crc_mulmod(p, sq)
        sq = _crc_mulmod(sq, sq)
        k >>= 1
    return _crc_mulmod(p, crc_a) ^ crc_b


def crc32_parallel(a, seg=SEG, threads=16):
    """zlib CRC32 of a large uint8 array, segments in threads, then combined."""
    a = np.ascontiguousarray(a)
    offs = list(range(0, a.size, seg)) or [0]
    with ThreadPoolExecutor(min(threads, len(offs))) as ex:
        parts = list(ex.map(lambda o: (zlib.crc32(memoryview(a[o:o + seg])), min(seg, a.size - o)), offs))
    crc = 0
    for c, n in parts:
        crc = crc32_combine(crc, c, n) if crc or n else c
    return crc


ENC_SEG = 8 << 20  # pigz-style encoder segment


def gzip_compress_segments(data, level, seg=ENC_SEG, threads=16):
    """One RFC 1952 member of `data` at `level`, encoded pigz-style in
    parallel: zlib raw deflate per `seg`-byte segment, each primed with the
    32 KiB before it (back-references cross the segment seams) and ended with
    a sync flush (no BFINAL, byte-aligned), the last one finished.  The
    concatenation is a single deflate stream, so the result is one member of
    stored, fixed and dynamic blocks however large `data` is."""
    a = np.ascontiguousarray(np.frombuffer(data, dtype=np.uint8) if not isinstance(data, np.ndarray) else data)
    offs = list(range(0, a.size, seg)) or [0]

    def one(i):
        o = offs[i]
        kw = {}
        if o:
            kw["zdict"] = bytes(a[max(0, o - 32768):o])
        co = zlib.compressobj(int(level), zlib.DEFLATED, -15, 9, zlib.Z_DEFAULT_STRATEGY, **kw)
        body = co.compress(memoryview(a[o:o + seg]))
        body += co.flush(zlib.Z_FINISH if i == len(offs) - 1 else zlib.Z_SYNC_FLUSH)
        return body

    with ThreadPoolExecutor(min(threads, len(offs))) as ex:
        bodies = list(ex.map(one, range(len(offs))))
    crc = crc32_parallel(a, seg, threads)
    head = b"\x1f\x8b\x08\x00\x00\x00\x00\x00\x00\xff"
    out = np.empty(len(head) + sum(len(b) for b in bodies) + 8, dtype=np.uint8)
    out[:len(head)] = np.frombuffer(head, np.uint8)
    pos = len(head)
    for b in bodies:
        out[pos:pos + len(b)] = np.frombuffer(b, np.uint8)
        pos += len(b)
    out[pos:] = np.frombuffer(struct.pack("<II", crc, a.size & 0xFFFFFFFF), np.uint8)
    return out


def mixed_member(n, seed, level, seg=SEG, threads=16):
    """Mixed-entropy single member: `seg`-byte segments cycling text / uniform
    random / byte runs / u32 deltas, encoded pigz-style (zlib) at `level`.
    Returns (raw, gz)."""
    raw = generate_mixed(n, seed, seg, threads=threads)
    return raw, gzip_compress_segments(raw, level, ENC_SEG, threads)


def trailer(gz):
    """(crc32, isize) from a member's 8-byte trailer."""
    crc, isize = struct.unpack("<II", bytes(gz[-8:]))
    return crc, isize


# ---------------------------------------------------------------- workloads

def text_member(n, seed, level=6):
    """Single-member shape: `n` bytes of the text generator, one member at `level`."""
    raw = generate("text", n, seed)
    return raw, gzip_compress(raw, level)


def batch_members(count, seed, min_size=1 << 10, max_size=64 << 20, median=64 << 10, sigma=1.5,
                  levels=(1, 6, 9), threads=8):
    """Batch shape: `count` independent members with lognormal sizes, content
    uniform over the four kinds and level uniform over `levels`.
    Returns list of (raw, gz, kind, level)."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(rng.normal(np.log(median), sigma, count)), min_size, max_size).astype(np.int64)
    kinds = rng.integers(0, 4, count)
    lv = rng.choice(np.array(levels), count)
    names = list(KINDS)

    def one(i):
        raw = generate(names[kinds[i]], int(sizes[i]), seed * 7919 + i)
        return raw, gzip_compress(raw, int(lv[i])), names[kinds[i]], int(lv[i])

    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, range(count)))


BGZF_EOF = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")


def sized_member_stream(raw, block, level=6, form="gz"):
    """Members of `block` input bytes each that carry their compressed size in
    FEXTRA.  form "gz": gzippy's parallel format, subfield "GZ" with the 4-byte
    whole-member size (src/compress/parallel.rs:470-499); "bgzf": subfield "BC"
    with the 2-byte size - 1 and the 28-byte empty end-of-file member
    (`block` <= 65280 there).  Bodies are zlib raw deflate at `level`."""
    raw = bytes(raw)
    parts = []
    for off in range(0, len(raw), block):
        chunk = raw[off:off + block]
        co = zlib.compressobj(int(level), zlib.DEFLATED, -15)
        body = co.compress(chunk) + co.flush()
        tail = struct.pack("<II", zlib.crc32(chunk), len(chunk))
        if form == "gz":
            size = 10 + 2 + 8 + len(body) + 8
            hdr = b"\x1f\x8b\x08\x04\0\0\0\0\0\xff" + struct.pack("<H", 8) + b"GZ" + struct.pack("<HI", 4, size)
        elif form == "bgzf":
            size = 10 + 2 + 6 + len(body) + 8
            if size > 65536:
                raise ValueError("BGZF member over 64 KiB")
            hdr = b"\x1f\x8b\x08\x04\0\0\0\0\0\xff" + struct.pack("<H", 6) + b"BC" + struct.pack("<HH", 2, size - 1)
        else:
            raise ValueError(form)
        parts.append(hdr + body + tail)
    if form == "bgzf":
        parts.append(BGZF_EOF)
    return b"".join(parts)
