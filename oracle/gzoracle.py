"""ctypes binding of the C oracle (``libgzoracle.so``) — test infrastructure.

Functions mirror the reference's CPU path; see ``gzoracle.h`` for the
file:line each one restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libgzoracle.so")

OK, INVALID_HEADER, INVALID_FORMAT, DECODE_FAILED, SIZE_MISMATCH, CRC_MISMATCH, OUTPUT_FULL = range(7)


class Block(ctypes.Structure):
    _fields_ = [("bit", ctypes.c_uint64), ("out_off", ctypes.c_uint64),
                ("btype", ctypes.c_uint32), ("bfinal", ctypes.c_uint32)]


class Result(ctypes.Structure):
    _fields_ = [("end_bit", ctypes.c_uint64), ("out_len", ctypes.c_uint64),
                ("status", ctypes.c_int32), ("bfinal_hit", ctypes.c_int32),
                ("n_blocks", ctypes.c_uint64)]


class Boundary(ctypes.Structure):
    _fields_ = [("bit_offset", ctypes.c_uint64), ("seek_bit", ctypes.c_uint64),
                ("kind", ctypes.c_uint32), ("hlit", ctypes.c_uint8), ("hdist", ctypes.c_uint8),
                ("hclen", ctypes.c_uint8), ("pad", ctypes.c_uint8)]


def build():
    """Compile libgzoracle.so from gzoracle.c (make, in place)."""
    subprocess.run(["make", "-s", "-C", _HERE, "libgzoracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.gzo_crc32.restype = ctypes.c_uint32
        L.gzo_crc32.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t]
        L.gzo_crc32_combine.restype = ctypes.c_uint32
        L.gzo_crc32_combine.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]
        L.gzo_x_pow_mod.restype = ctypes.c_uint32
        L.gzo_x_pow_mod.argtypes = [ctypes.c_uint64]
        L.gzo_mulmod.restype = ctypes.c_uint32
        L.gzo_mulmod.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.gzo_parse_header.restype = ctypes.c_int
        L.gzo_parse_header.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.gzo_inflate.restype = ctypes.c_int
        L.gzo_inflate.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                  ctypes.POINTER(Result)]
        L.gzo_resolve_markers.restype = ctypes.c_int
        L.gzo_resolve_markers.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
        L.gzo_is_dynamic_candidate.restype = ctypes.c_int
        L.gzo_is_dynamic_candidate.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64]
        L.gzo_find_blocks.restype = ctypes.c_size_t
        L.gzo_find_blocks.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64,
                                      ctypes.c_void_p, ctypes.c_size_t]
        L.gzo_gunzip.restype = ctypes.c_int
        L.gzo_gunzip.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        L.gzo_gunzip_parallel.restype = ctypes.c_int
        L.gzo_gunzip_parallel.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t), ctypes.c_int, ctypes.c_size_t]
        del u8p
        _lib = L
    return _lib


def _buf(data):
    a = np.frombuffer(data, dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    return np.ascontiguousarray(a, dtype=np.uint8)


def crc32(data, crc=0):
    a = _buf(data)
    return lib().gzo_crc32(crc, a.ctypes.data, a.size)


def crc32_combine(c1, c2, len2):
    return lib().gzo_crc32_combine(c1, c2, len2)


def parse_header(data):
    a = _buf(data)
    n = ctypes.c_size_t()
    st = lib().gzo_parse_header(a.ctypes.data, a.size, ctypes.byref(n))
    return st, n.value


def inflate(data, start_bit=0, stop_bit=2**64 - 1, marker_mode=False, window=None, out_cap=None,
            blocks_cap=0):
    """Decode raw DEFLATE from start_bit; returns (Result, out array, blocks list)."""
    a = _buf(data)
    if out_cap is None:
        out_cap = max(1 << 16, a.size * 1032 + 65536)
    out = np.zeros(out_cap, dtype=np.uint16 if marker_mode else np.uint8)
    w = None
    if window is not None:
        w = _buf(window)
        assert w.size == 32768
    blk = (Block * max(blocks_cap, 1))()
    res = Result()
    lib().gzo_inflate(a.ctypes.data, a.size, start_bit, stop_bit, 1 if marker_mode else 0,
                      None if w is None else w.ctypes.data,
                      None if marker_mode else out.ctypes.data,
                      out.ctypes.data if marker_mode else None,
                      out_cap, ctypes.addressof(blk) if blocks_cap else None, blocks_cap,
                      ctypes.byref(res))
    nb = min(res.n_blocks, blocks_cap)
    blocks = [(blk[i].bit, blk[i].out_off, blk[i].btype, blk[i].bfinal) for i in range(nb)]
    return res, out[: res.out_len], blocks


def resolve_markers(markers, window):
    m = np.ascontiguousarray(markers, dtype=np.uint16)
    w = _buf(window)
    out = np.zeros(m.size, dtype=np.uint8)
    st = lib().gzo_resolve_markers(m.ctypes.data, m.size, w.ctypes.data, out.ctypes.data)
    return st, out


def is_dynamic_candidate(data, bit):
    a = _buf(data)
    return bool(lib().gzo_is_dynamic_candidate(a.ctypes.data, a.size, bit))


def find_blocks(data, start_bit=0, end_bit=None, cap=1 << 20):
    """Sorted list of (bit_offset, seek_bit, kind) candidates in [start_bit, end_bit)."""
    a = _buf(data)
    if end_bit is None:
        end_bit = a.size * 8
    out = (Boundary * cap)()
    n = lib().gzo_find_blocks(a.ctypes.data, a.size, start_bit, end_bit, ctypes.addressof(out), cap)
    if n > cap:
        return find_blocks(data, start_bit, end_bit, cap=n)
    return [(out[i].bit_offset, out[i].seek_bit, out[i].kind) for i in range(n)]


def gunzip(data, out_cap=None):
    """Serial whole-stream decode; returns (status, bytes, members)."""
    a = _buf(data)
    if out_cap is None:
        out_cap = max(1 << 16, a.size * 1032 + 65536)
    out = np.zeros(out_cap, dtype=np.uint8)
    n = ctypes.c_size_t()
    m = ctypes.c_size_t()
    st = lib().gzo_gunzip(a.ctypes.data, a.size, out.ctypes.data, out_cap, ctypes.byref(n), ctypes.byref(m))
    return st, out[: n.value], m.value


def gunzip_parallel(data, threads=4, chunk_bytes=512 << 10, out_cap=None):
    """Threaded speculative decode of one member (gzpar.c); returns
    (status, bytes)."""
    a = _buf(data)
    if out_cap is None:
        out_cap = max(1 << 16, int(a[-4:].view(np.uint32)[0]) + 65536 if a.size >= 18 else 1 << 16)
    out = np.empty(out_cap, dtype=np.uint8)
    n = ctypes.c_size_t()
    st = lib().gzo_gunzip_parallel(a.ctypes.data, a.size, out.ctypes.data, out_cap, ctypes.byref(n), int(threads),
                                   int(chunk_bytes))
    return st, out[: n.value]
