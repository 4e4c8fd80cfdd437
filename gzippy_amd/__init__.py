"""gzippy_amd — MI355X (gfx950) parallel gzip decompression.

Host-side mirror of gzippy's public decompression API (src/lib.rs:166-227),
backed by the HIP library ``libgzgpu.so`` through its C ABI:

    decompress(data)                              lib.rs:176
    decompress_with_threads(data, threads)        lib.rs:189
    decompress_to_writer(data, writer)            lib.rs:204
    decompress_to_writer_with_threads(...)        lib.rs:221

Semantics kept from the reference: non-gzip input returns empty output
(lib.rs:50-52); multi-member input concatenates; bytes after a valid member
that do not start a member end the walk cleanly; corruption raises
``GzippyError`` (the ``GzippyError::Decompression`` arm, error.rs:17) — there
is no fallback to another decoder.  ``threads`` is accepted for signature
parity; the GPU picks its own parallelism.

Device-resident entry points (inputs/outputs as torch CUDA tensors):
``decompress_device``, ``decompress_batch``, ``find_blocks``, ``crc32``,
``inflate_piece``.
"""
import ctypes
import threading

import numpy as np

from . import _native
from ._native import GzgpuError, Stats, Tuning  # noqa: F401

__all__ = [
    "GzippyError", "decompress", "decompress_with_threads", "decompress_to_writer",
    "decompress_to_writer_with_threads", "decompress_device", "decompress_batch", "find_blocks", "crc32",
    "inflate_piece", "context", "Server", "autotune", "BatchGraph", "build_index", "seek_decompress",
    "decompress_members_plain",
]


class GzippyError(Exception):
    """Mirror of gzippy's GzippyError (src/error.rs:7-46).  ``kind`` is the
    ParallelError / device status name (single_member.rs:1384-1413)."""

    def __init__(self, message, kind="Decompression", status=None):
        super().__init__(f"Decompression error: {message}")
        self.kind = kind
        self.status = status


_tls = threading.local()


def context(device=0):
    """The calling thread's device context (created on first use)."""
    ctxs = getattr(_tls, "ctxs", None)
    if ctxs is None:
        ctxs = _tls.ctxs = {}
    c = ctxs.get(device)
    if c is None:
        c = ctxs[device] = _native.Context(device)
    return c


def _raise(ctx, st):
    e = ctx.error(st)
    raise GzippyError(str(e), kind=e.kind, status=st)


def _as_bytes_view(data):
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8)
    return np.frombuffer(memoryview(data), dtype=np.uint8)


def decompress(data, stats=None):
    """Decompress a gzip stream; returns bytes (lib.rs:176)."""
    return decompress_with_threads(data, 0, stats=stats)


def decompress_with_threads(data, threads, stats=None):
    """lib.rs:189.  `threads` is accepted for API parity; unused."""
    ctx = context()
    a = _as_bytes_view(data)
    n = ctypes.c_size_t()
    st_obj = Stats()
    cap = max(1 << 20, 4 * a.size)
    if a.size >= 8 and a.size >= 2 and a[0] == 0x1F and a[1] == 0x8B:
        cap = max(cap, int(a[-4:].view(np.uint32)[0]) + 1)
    for _ in range(16):
        out = np.empty(cap, dtype=np.uint8)
        st = ctx.lib.gzgpu_decompress(ctx.h, a.ctypes.data, a.size, out.ctypes.data, cap, ctypes.byref(n),
                                      ctypes.byref(st_obj))
        if st == _native.OUTPUT_TOO_SMALL and n.value > cap:
            cap = n.value
            continue
        if st != _native.OK:
            _raise(ctx, st)
        if stats is not None:
            stats.update(st_obj.as_dict())
        return out[: n.value].tobytes()
    _raise(ctx, _native.OUTPUT_TOO_SMALL)


def decompress_to_writer(data, writer, stats=None):
    """Stream the decompressed bytes into `writer` (anything with .write);
    returns the byte count (lib.rs:204)."""
    return decompress_to_writer_with_threads(data, writer, 0, stats=stats)


def decompress_to_writer_with_threads(data, writer, threads, stats=None):
    ctx = context()
    a = _as_bytes_view(data)
    failure = []

    @_native.WRITE_FN
    def _cb(user, ptr, n):
        try:
            writer.write(ctypes.string_at(ptr, n))
            return 0
        except Exception as e:  # propagate as GZGPU_IO (ParallelError::Io)
            failure.append(e)
            return 1

    total = ctypes.c_uint64()
    st_obj = Stats()
    st = ctx.lib.gzgpu_decompress_to_writer(ctx.h, a.ctypes.data, a.size, _cb, None, ctypes.byref(total),
                                            ctypes.byref(st_obj))
    if st != _native.OK:
        if failure:
            raise GzippyError(f"I/O error: {failure[0]}", kind="Io", status=st) from failure[0]
        _raise(ctx, st)
    if stats is not None:
        stats.update(st_obj.as_dict())
    return total.value


def build_index(data, interval_bytes=1 << 20):
    """Seek index of a single-member gzip stream (index.rs:54 build_index,
    :145 serialize_index): bytes in index.rs's version-1 layout."""
    ctx = context()
    a = _as_bytes_view(data)
    need = ctypes.c_size_t()
    st = ctx.lib.gzgpu_index_build(ctx.h, a.ctypes.data, a.size, int(interval_bytes), None, 0, ctypes.byref(need))
    if st not in (_native.OK, _native.OUTPUT_TOO_SMALL):
        _raise(ctx, st)
    buf = np.empty(need.value, dtype=np.uint8)
    st = ctx.lib.gzgpu_index_build(ctx.h, a.ctypes.data, a.size, int(interval_bytes), buf.ctypes.data, buf.size,
                                   ctypes.byref(need))
    if st != _native.OK:
        _raise(ctx, st)
    return buf[: need.value].tobytes()


def seek_decompress(data, index, offset, max_bytes=None):
    """Decompressed bytes [offset, offset + max_bytes) of `data` through its
    index (index.rs:344 seek_decompress); max_bytes None: to the end."""
    ctx = context()
    a = _as_bytes_view(data)
    ix = _as_bytes_view(index)
    total = int(ix[10:18].view(np.uint64)[0]) if ix.size >= 18 else 0
    if max_bytes is None:
        max_bytes = max(0, total - offset)
    cap = max(1, min(int(max_bytes), max(0, total - int(offset))))
    out = np.empty(cap, dtype=np.uint8)
    n = ctypes.c_size_t()
    st = ctx.lib.gzgpu_index_seek(ctx.h, a.ctypes.data, a.size, ix.ctypes.data, ix.size, int(offset), int(max_bytes),
                                  out.ctypes.data, out.size, ctypes.byref(n))
    if st != _native.OK:
        _raise(ctx, st)
    return out[: n.value].tobytes()


# ---------------------------------------------------------------- device-resident


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _check(t, name, dtypes=("uint8",), length=None):
    """Device tensors handed to the kernels: CUDA, contiguous, of the expected
    element type, and at least `length` bytes long (the kernels read and write
    exactly what the lengths say)."""
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise GzippyError(f"{name} must be a CUDA tensor", kind="InvalidArgument")
    if str(t.dtype).replace("torch.", "") not in dtypes:
        raise GzippyError(f"{name} must be {'/'.join(dtypes)}, got {t.dtype}", kind="InvalidArgument")
    if not t.is_contiguous():
        raise GzippyError(f"{name} must be contiguous", kind="InvalidArgument")
    if length is not None and (length < 0 or length > t.numel() * t.element_size()):
        raise GzippyError(f"{name} holds {t.numel() * t.element_size()} bytes, {length} asked",
                          kind="InvalidArgument")


def _check_ranges(name, off, ln, t):
    total = t.numel()
    for o, n in zip(off, ln):
        if int(o) + int(n) > total:
            raise GzippyError(f"{name} range [{int(o)}, {int(o) + int(n)}) past the tensor's {total} bytes",
                              kind="InvalidArgument")


def _sync_torch(t):
    """The library runs on its own HIP stream: drain torch's current stream so
    work it queued on these tensors (copies, fills) is complete first."""
    import torch
    torch.cuda.current_stream(t.device).synchronize()


def decompress_device(d_in, d_out, in_len=None, ctx=None):
    """Decode one gzip member from device tensor `d_in` into `d_out`
    (uint8 CUDA tensors).  Returns (out_len, crc32, stats dict)."""
    L = d_in.numel() if in_len is None else in_len
    _check(d_in, "d_in", length=L)
    _check(d_out, "d_out")
    ctx = ctx or context(d_in.device.index or 0)
    _sync_torch(d_in)
    n = ctypes.c_size_t()
    crc = ctypes.c_uint32()
    st_obj = Stats()
    st = ctx.lib.gzgpu_decompress_device(ctx.h, _ptr(d_in), L, _ptr(d_out), d_out.numel(), ctypes.byref(n),
                                         ctypes.byref(crc), ctypes.byref(st_obj))
    if st != _native.OK:
        _raise(ctx, st)
    return n.value, crc.value, st_obj.as_dict()


def decompress_batch(d_in, in_off, in_len, d_out, out_off, out_cap, ctx=None, raise_on_error=True):
    """Decode n independent members in one persistent launch.
    Offsets/lengths are host integer sequences.  Returns (status[], out_len[],
    crc[], stats dict) as numpy arrays."""
    _check(d_in, "d_in")
    _check(d_out, "d_out")
    _check_ranges("input", in_off, in_len, d_in)
    _check_ranges("output", out_off, out_cap, d_out)
    ctx = ctx or context(d_in.device.index or 0)
    _sync_torch(d_in)
    n = len(in_off)
    a = [np.ascontiguousarray(np.asarray(x, dtype=np.uint64)) for x in (in_off, in_len, out_off, out_cap)]
    status = np.zeros(n, dtype=np.int32)
    olen = np.zeros(n, dtype=np.uint64)
    crcs = np.zeros(n, dtype=np.uint32)
    st_obj = Stats()
    P64 = ctypes.POINTER(ctypes.c_uint64)
    st = ctx.lib.gzgpu_decompress_batch(
        ctx.h, _ptr(d_in), a[0].ctypes.data_as(P64), a[1].ctypes.data_as(P64), _ptr(d_out),
        a[2].ctypes.data_as(P64), a[3].ctypes.data_as(P64), n,
        status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), olen.ctypes.data_as(P64),
        crcs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(st_obj))
    if st != _native.OK and raise_on_error:
        _raise(ctx, st)
    return status, olen, crcs, st_obj.as_dict()


def decompress_members_plain(d_in, in_off, in_len, d_out, out_off, out_cap, ctx=None, raise_on_error=True):
    """The plain per-member form: one inflate launch per member, back to back,
    no host synchronisation in between.  Returns (status[], out_len[], crc[],
    stats dict)."""
    _check(d_in, "d_in")
    _check(d_out, "d_out")
    _check_ranges("input", in_off, in_len, d_in)
    _check_ranges("output", out_off, out_cap, d_out)
    ctx = ctx or context(d_in.device.index or 0)
    _sync_torch(d_in)
    n = len(in_off)
    a = [np.ascontiguousarray(np.asarray(x, dtype=np.uint64)) for x in (in_off, in_len, out_off, out_cap)]
    status = np.zeros(n, dtype=np.int32)
    olen = np.zeros(n, dtype=np.uint64)
    crcs = np.zeros(n, dtype=np.uint32)
    st_obj = Stats()
    P64 = ctypes.POINTER(ctypes.c_uint64)
    st = ctx.lib.gzgpu_decompress_members_plain(
        ctx.h, _ptr(d_in), a[0].ctypes.data_as(P64), a[1].ctypes.data_as(P64), _ptr(d_out),
        a[2].ctypes.data_as(P64), a[3].ctypes.data_as(P64), n,
        status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), olen.ctypes.data_as(P64),
        crcs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(st_obj))
    if st != _native.OK and raise_on_error:
        _raise(ctx, st)
    return status, olen, crcs, st_obj.as_dict()


class BatchGraph:
    """A fixed batch layout captured as one hipGraph (gzgpu_batch_graph_*):
    header parse, block finder and piece plan for the members worth splitting,
    the persistent inflate kernel fed by device-side queues, link checks and
    re-decodes, window chain, marker resolution, CRC32 and trailer checks, all
    on the device.  launch() replays it; the contents of d_in may change
    between launches, the layout may not.  stats: pieces (members plus split
    pieces), redecoded, candidates (block-finder hits)."""

    def __init__(self, d_in, in_off, in_len, d_out, out_off, out_cap, ctx=None):
        _check(d_in, "d_in")
        _check(d_out, "d_out")
        _check_ranges("input", in_off, in_len, d_in)
        _check_ranges("output", out_off, out_cap, d_out)
        self.ctx = ctx or context(d_in.device.index or 0)
        self._keep = (d_in, d_out)
        _sync_torch(d_in)
        self.n = len(in_off)
        a = [np.ascontiguousarray(np.asarray(x, dtype=np.uint64)) for x in (in_off, in_len, out_off, out_cap)]
        P64 = ctypes.POINTER(ctypes.c_uint64)
        h = ctypes.c_void_p()
        st = self.ctx.lib.gzgpu_batch_graph_create(
            self.ctx.h, _ptr(d_in), a[0].ctypes.data_as(P64), a[1].ctypes.data_as(P64), _ptr(d_out),
            a[2].ctypes.data_as(P64), a[3].ctypes.data_as(P64), self.n, ctypes.byref(h))
        if st != _native.OK:
            _raise(self.ctx, st)
        self.h = h

    def launch(self, raise_on_error=True):
        """Replay; returns (status[], out_len[], crc[], stats dict)."""
        _sync_torch(self._keep[0])
        status = np.zeros(self.n, dtype=np.int32)
        olen = np.zeros(self.n, dtype=np.uint64)
        crcs = np.zeros(self.n, dtype=np.uint32)
        st_obj = Stats()
        st = self.ctx.lib.gzgpu_batch_graph_launch(
            self.h, status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            olen.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), crcs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            ctypes.byref(st_obj))
        if st != _native.OK and raise_on_error:
            _raise(self.ctx, st)
        return status, olen, crcs, st_obj.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.gzgpu_batch_graph_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(d_in, in_off, in_len, d_out, out_off, out_cap, budget_ms=2000, ctx=None):
    """Pick the fastest tuning for these members (device time) and keep it in
    the context; returns (Tuning, ms)."""
    _check(d_in, "d_in")
    _check(d_out, "d_out")
    _check_ranges("input", in_off, in_len, d_in)
    _check_ranges("output", out_off, out_cap, d_out)
    ctx = ctx or context(d_in.device.index or 0)
    _sync_torch(d_in)
    a = [np.ascontiguousarray(np.asarray(x, dtype=np.uint64)) for x in (in_off, in_len, out_off, out_cap)]
    P64 = ctypes.POINTER(ctypes.c_uint64)
    best = Tuning()
    ms = ctypes.c_double()
    st = ctx.lib.gzgpu_autotune(ctx.h, _ptr(d_in), a[0].ctypes.data_as(P64), a[1].ctypes.data_as(P64), _ptr(d_out),
                                a[2].ctypes.data_as(P64), a[3].ctypes.data_as(P64), len(in_off), int(budget_ms),
                                ctypes.byref(best), ctypes.byref(ms))
    if st != _native.OK:
        _raise(ctx, st)
    return best, ms.value


def find_blocks(d_in, start_bit=0, end_bit=None, in_len=None, ctx=None):
    """Block-finder hits in [start_bit, end_bit) of a device buffer, sorted:
    list of (bit_offset, seek_bit, kind)."""
    L = d_in.numel() if in_len is None else in_len
    _check(d_in, "d_in", length=L)
    ctx = ctx or context(d_in.device.index or 0)
    _sync_torch(d_in)
    if end_bit is None:
        end_bit = L * 8
    cnt = ctypes.c_size_t()
    cap = 1 << 16
    while True:
        buf = (_native.Boundary * cap)()
        st = ctx.lib.gzgpu_find_blocks(ctx.h, _ptr(d_in), L, start_bit, end_bit, buf, cap, ctypes.byref(cnt))
        if st != _native.OK:
            _raise(ctx, st)
        if cnt.value <= cap:
            return [(buf[i].bit_offset, buf[i].seek_bit, buf[i].kind) for i in range(cnt.value)]
        cap = cnt.value


def crc32(d_data, length=None, ctx=None):
    """gzip CRC32 of a device buffer."""
    L = d_data.numel() * d_data.element_size() if length is None else length
    _check(d_data, "d_data", dtypes=("uint8", "int8", "int16", "uint16", "int32", "uint32", "int64", "uint64"),
           length=L)
    ctx = ctx or context(d_data.device.index or 0)
    _sync_torch(d_data)
    c = ctypes.c_uint32()
    st = ctx.lib.gzgpu_crc32(ctx.h, _ptr(d_data), L, ctypes.byref(c))
    if st != _native.OK:
        _raise(ctx, st)
    return c.value


def inflate_piece(d_in, start_bit, stop_bit, d_out, window=None, in_len=None, ctx=None):
    """One wavefront decodes raw DEFLATE from start_bit with the stop rule.
    window=None: marker mode, d_out is int16/uint16-sized entries; window =
    32768 host bytes: clean decode into uint8 d_out.  Returns PieceResult."""
    L = d_in.numel() if in_len is None else in_len
    _check(d_in, "d_in", length=L)
    _check(d_out, "d_out", dtypes=("uint8",) if window is not None else ("int16", "uint16"))
    ctx = ctx or context(d_in.device.index or 0)
    _sync_torch(d_in)
    res = _native.PieceResult()
    w = None
    if window is not None:
        w = np.ascontiguousarray(np.frombuffer(bytes(window), dtype=np.uint8))
        assert w.size == 32768
    st = ctx.lib.gzgpu_inflate_piece(ctx.h, _ptr(d_in), L, start_bit, stop_bit,
                                     None if w is None else w.ctypes.data, _ptr(d_out), d_out.numel(),
                                     ctypes.byref(res))
    if st != _native.OK:
        _raise(ctx, st)
    return res


class Server:
    """Continuous-batching decompression server on one GPU (the C ABI's
    gzgpu_server).  submit() copies the request and returns a ticket at once;
    requests that arrive while others run join the next device batch.  Each
    result equals decompress() of that request alone."""

    def __init__(self, device=0, hbm_budget_bytes=0, max_batch_requests=0, workers=0):
        self.lib = _native.load_library()
        cfg = _native.ServerConfig(hbm_budget_bytes, max_batch_requests, workers)
        h = ctypes.c_void_p()
        st = self.lib.gzgpu_server_create(device, ctypes.byref(cfg), ctypes.byref(h))
        if st != _native.OK:
            raise GzippyError("cannot start the server", kind=_native.STATUS.get(st, str(st)), status=st)
        self.h = h

    def close(self):
        """Drain queued requests and stop the workers."""
        if getattr(self, "h", None):
            self.lib.gzgpu_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, data):
        a = _as_bytes_view(data)
        t = ctypes.c_uint64()
        st = self.lib.gzgpu_submit(self.h, a.ctypes.data if a.size else None, a.size, ctypes.byref(t))
        if st != _native.OK:
            raise GzippyError("submit failed", kind=_native.STATUS.get(st, str(st)), status=st)
        return t.value

    def poll(self, ticket, timeout_ms=-1):
        """Request info dict once finished (waits up to timeout_ms; < 0: wait),
        else None."""
        info = _native.RequestInfo()
        st = self.lib.gzgpu_poll(self.h, ticket, int(timeout_ms), ctypes.byref(info))
        if st == _native.PENDING:
            return None
        if st != _native.OK:
            raise GzippyError("unknown ticket", kind=_native.STATUS.get(st, str(st)), status=st)
        return {"status": info.status, "members": info.members, "out_len": info.out_len,
                "queue_ms": info.queue_ms, "latency_ms": info.latency_ms, "batch": info.batch}

    def fetch(self, ticket, info=None):
        """The decoded bytes of a finished request (releases the ticket);
        raises GzippyError when the request failed."""
        info = info or self.poll(ticket, -1)
        out = np.empty(max(int(info["out_len"]), 1), dtype=np.uint8)
        n = ctypes.c_size_t()
        st = self.lib.gzgpu_fetch(self.h, ticket, out.ctypes.data, out.size, ctypes.byref(n))
        if st != _native.OK:
            raise GzippyError(f"request failed: {_native.STATUS.get(st, st)}", kind=_native.STATUS.get(st, str(st)),
                              status=st)
        return out[: n.value].tobytes()

    def fetch_into(self, ticket, out):
        """Copy a finished request's bytes into the uint8 numpy array `out`;
        returns the byte count (releases the ticket)."""
        n = ctypes.c_size_t()
        st = self.lib.gzgpu_fetch(self.h, ticket, out.ctypes.data, out.size, ctypes.byref(n))
        if st != _native.OK:
            raise GzippyError(f"request failed: {_native.STATUS.get(st, st)}", kind=_native.STATUS.get(st, str(st)),
                              status=st)
        return n.value

    def result(self, ticket):
        return self.fetch(ticket, self.poll(ticket, -1))

    def stats(self):
        s = _native.ServerStats()
        self.lib.gzgpu_server_get_stats(self.h, ctypes.byref(s))
        return s.as_dict()
