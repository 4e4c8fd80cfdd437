"""ctypes binding of ``libgzgpu.so`` (C ABI declared in ``include/gzgpu.h``).

The library is loaded from this package directory only; if it is missing or
no gfx950 device is present, every entry point raises — there is no CPU
fallback.
"""
import ctypes
import os
import threading

from . import build as _build

STATUS = {
    0: "OK", 1: "InvalidHeader", 2: "InvalidGzipFormat", 3: "DecodeFailed", 4: "SizeMismatch",
    5: "CrcMismatch", 6: "Io", 7: "GpuUnavailable", 8: "OutOfDeviceMemory", 9: "OutputTooSmall",
    10: "InvalidArgument", 11: "Pending",
}
OK, INVALID_HEADER, INVALID_FORMAT, DECODE_FAILED, SIZE_MISMATCH, CRC_MISMATCH, IO, GPU_UNAVAILABLE, \
    OUT_OF_DEVICE_MEMORY, OUTPUT_TOO_SMALL, INVALID_ARGUMENT = range(11)

EXPORTS = [
    "gzgpu_ctx_create", "gzgpu_ctx_destroy", "gzgpu_set_tuning", "gzgpu_get_tuning", "gzgpu_last_error",
    "gzgpu_status_string", "gzgpu_decompress", "gzgpu_decompress_to_writer", "gzgpu_decompress_device",
    "gzgpu_decompress_batch", "gzgpu_find_blocks", "gzgpu_crc32", "gzgpu_inflate_piece",
    "gzgpu_server_create", "gzgpu_server_destroy", "gzgpu_submit", "gzgpu_poll", "gzgpu_fetch",
    "gzgpu_server_get_stats", "gzgpu_autotune", "gzgpu_batch_graph_create", "gzgpu_batch_graph_launch",
    "gzgpu_batch_graph_destroy", "gzgpu_decompress_device_noalloc", "gzgpu_index_build", "gzgpu_index_seek",
    "gzgpu_decompress_members_plain",
]
PENDING = 11


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_batch_requests", ctypes.c_uint32),
                ("workers", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("members", ctypes.c_uint32), ("out_len", ctypes.c_uint64),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("batch", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests_done", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("bytes_in", ctypes.c_uint64),
                ("bytes_out", ctypes.c_uint64), ("fallbacks", ctypes.c_uint64), ("hbm_peak_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("chunk_bytes", ctypes.c_uint64), ("pieces_per_chunk", ctypes.c_uint32),
                ("waves_per_cu", ctypes.c_uint32), ("batch_split_bytes", ctypes.c_uint64),
                ("miss_every", ctypes.c_uint32), ("lut_bits", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_find", ctypes.c_double), ("ms_inflate", ctypes.c_double),
                ("ms_resolve", ctypes.c_double), ("ms_crc", ctypes.c_double),
                ("compressed_bytes", ctypes.c_uint64), ("decompressed_bytes", ctypes.c_uint64),
                ("pieces", ctypes.c_uint32), ("redecoded", ctypes.c_uint32), ("candidates", ctypes.c_uint32),
                ("launches", ctypes.c_uint32), ("ms_inflate_kernel", ctypes.c_double),
                ("inflate_launches", ctypes.c_uint32), ("pad_", ctypes.c_uint32),
                ("ms_inflate_marker", ctypes.c_double), ("ms_inflate_clean", ctypes.c_double),
                ("clean_in_bytes", ctypes.c_uint64), ("clean_out_bytes", ctypes.c_uint64),
                ("marker_in_bytes", ctypes.c_uint64), ("marker_out_bytes", ctypes.c_uint64),
                ("midblock_pieces", ctypes.c_uint32), ("midblock_linked", ctypes.c_uint32),
                ("stored_members", ctypes.c_uint32), ("stored_blocks", ctypes.c_uint32),
                ("final_block_pieces", ctypes.c_uint32), ("stored_runs_skipped", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Boundary(ctypes.Structure):
    _fields_ = [("bit_offset", ctypes.c_uint64), ("seek_bit", ctypes.c_uint64), ("kind", ctypes.c_uint32),
                ("info", ctypes.c_uint32)]


class PieceResult(ctypes.Structure):
    _fields_ = [("end_bit", ctypes.c_uint64), ("out_len", ctypes.c_uint64), ("status", ctypes.c_int32),
                ("bfinal_hit", ctypes.c_int32)]


WRITE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t)

_lib = None
_lock = threading.Lock()


def lib_path():
    return _build.gpu_lib_path()


def load_library():
    """dlopen libgzgpu.so and declare every exported signature (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"gzippy_amd: HIP library {path} is not built (run __graft_entry__.build())")
        # torch bundles its own libamdhip64.so.7 (same soname as /opt/rocm's).
        # Load torch first so the process holds exactly one HIP runtime: ours
        # then binds to it instead of pulling in a second HSA runtime, which
        # leaves torch without devices.
        import torch  # noqa: F401
        L = ctypes.CDLL(path)
        vp, sz, u64, u32, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
        P = ctypes.POINTER
        L.gzgpu_ctx_create.argtypes = [ctypes.c_int, P(vp)]
        L.gzgpu_ctx_create.restype = ctypes.c_int
        L.gzgpu_ctx_destroy.argtypes = [vp]
        L.gzgpu_ctx_destroy.restype = None
        L.gzgpu_set_tuning.argtypes = [vp, P(Tuning)]
        L.gzgpu_get_tuning.argtypes = [vp, P(Tuning)]
        L.gzgpu_last_error.argtypes = [vp]
        L.gzgpu_last_error.restype = ctypes.c_char_p
        L.gzgpu_status_string.argtypes = [ctypes.c_int]
        L.gzgpu_status_string.restype = ctypes.c_char_p
        L.gzgpu_decompress.argtypes = [vp, vp, sz, vp, sz, P(sz), P(Stats)]
        L.gzgpu_decompress_to_writer.argtypes = [vp, vp, sz, WRITE_FN, vp, P(u64), P(Stats)]
        L.gzgpu_decompress_device.argtypes = [vp, vp, sz, vp, sz, P(sz), P(u32), P(Stats)]
        L.gzgpu_decompress_device_noalloc.argtypes = [vp, vp, sz, vp, sz, P(sz), P(u32), P(Stats)]
        L.gzgpu_index_build.argtypes = [vp, vp, sz, u64, vp, sz, P(sz)]
        L.gzgpu_decompress_members_plain.argtypes = [vp, vp, P(u64), P(u64), vp, P(u64), P(u64), sz, P(i32), P(u64),
                                                     P(u32), P(Stats)]
        L.gzgpu_index_seek.argtypes = [vp, vp, sz, vp, sz, u64, u64, vp, sz, P(sz)]
        L.gzgpu_decompress_batch.argtypes = [vp, vp, P(u64), P(u64), vp, P(u64), P(u64), sz, P(i32), P(u64),
                                             P(u32), P(Stats)]
        L.gzgpu_find_blocks.argtypes = [vp, vp, sz, u64, u64, P(Boundary), sz, P(sz)]
        L.gzgpu_crc32.argtypes = [vp, vp, sz, P(u32)]
        L.gzgpu_inflate_piece.argtypes = [vp, vp, sz, u64, u64, vp, vp, sz, P(PieceResult)]
        L.gzgpu_server_create.argtypes = [ctypes.c_int, P(ServerConfig), P(vp)]
        L.gzgpu_server_destroy.argtypes = [vp]
        L.gzgpu_server_destroy.restype = None
        L.gzgpu_submit.argtypes = [vp, vp, sz, P(u64)]
        L.gzgpu_poll.argtypes = [vp, u64, ctypes.c_int32, P(RequestInfo)]
        L.gzgpu_fetch.argtypes = [vp, u64, vp, sz, P(sz)]
        L.gzgpu_server_get_stats.argtypes = [vp, P(ServerStats)]
        L.gzgpu_autotune.argtypes = [vp, vp, P(u64), P(u64), vp, P(u64), P(u64), sz, ctypes.c_uint32, P(Tuning),
                                     P(ctypes.c_double)]
        L.gzgpu_batch_graph_create.restype = ctypes.c_int
        L.gzgpu_batch_graph_create.argtypes = [vp, vp, P(u64), P(u64), vp, P(u64), P(u64), sz, P(vp)]
        L.gzgpu_batch_graph_launch.restype = ctypes.c_int
        L.gzgpu_batch_graph_launch.argtypes = [vp, P(ctypes.c_int32), P(u64), P(ctypes.c_uint32), P(Stats)]
        L.gzgpu_batch_graph_destroy.restype = None
        L.gzgpu_batch_graph_destroy.argtypes = [vp]
        for name in EXPORTS:
            getattr(L, name)  # every declared symbol must resolve
        _lib = L
        return L


class GzgpuError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        self.kind = STATUS.get(status, str(status))
        super().__init__(f"{self.kind}: {message}")


class Context:
    """One HIP device context (stream + grown device buffers).  Not thread-safe:
    use one per thread."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        st = self.lib.gzgpu_ctx_create(device, ctypes.byref(h))
        if st != OK:
            raise GzgpuError(st, "cannot create a gfx950 context on device %d" % device)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.gzgpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self, st):
        msg = self.lib.gzgpu_last_error(self.h)
        return GzgpuError(st, msg.decode() if msg else "")

    def tuning(self):
        t = Tuning()
        self.lib.gzgpu_get_tuning(self.h, ctypes.byref(t))
        return t

    def set_tuning(self, chunk_bytes=0, pieces_per_chunk=0, waves_per_cu=0, batch_split_bytes=0, miss_every=0,
                   lut_bits=0):
        t = Tuning(chunk_bytes, pieces_per_chunk, waves_per_cu, batch_split_bytes, miss_every, lut_bits)
        if self.lib.gzgpu_set_tuning(self.h, ctypes.byref(t)) != OK:
            raise ValueError(self.lib.gzgpu_last_error(self.h).decode())
