// src/decompress/parallel/gpu_sm.rs — a new module in gzippy (not compiled in
// this repository).  The ParallelSM contract of single_member::
// decompress_parallel (single_member.rs:987) on one MI355X through gzgpu-sys.
use std::io::Write;

use crate::decompress::parallel::single_member::ParallelError;

/// Status codes 1..=6 are ParallelError's variants (single_member.rs:1384-1413);
/// 7..=11 are device conditions, reported as DecodeFailed with the message.
fn to_parallel_error(e: gzgpu_sys::GpuError, io: Option<std::io::Error>) -> ParallelError {
    match e.status {
        gzgpu_sys::GZGPU_INVALID_HEADER => ParallelError::InvalidHeader,
        gzgpu_sys::GZGPU_INVALID_FORMAT => ParallelError::InvalidGzipFormat,
        gzgpu_sys::GZGPU_DECODE_FAILED => ParallelError::DecodeFailed(e.message),
        gzgpu_sys::GZGPU_SIZE_MISMATCH => ParallelError::SizeMismatch,
        gzgpu_sys::GZGPU_CRC_MISMATCH => ParallelError::CrcMismatch,
        gzgpu_sys::GZGPU_IO => ParallelError::Io(io.unwrap_or_else(|| std::io::Error::other(e.message))),
        _ => ParallelError::DecodeFailed(format!("gpu: {e}")),
    }
}

/// Same contract as decompress_parallel: in-order output to `writer`, CRC32
/// and ISIZE checked against the trailer, partial output allowed before a
/// terminal CRC/size error (single_member.rs:17-22).  `W: Write` only, as in
/// decompress_single_member_for (decompress/mod.rs:629): the writer is used
/// from the calling thread inside the callback.
pub fn decompress_parallel_gpu<W: Write>(data: &[u8], writer: &mut W) -> Result<u64, ParallelError> {
    gzgpu_sys::decompress_to_writer(data, writer).map_err(|(e, io)| to_parallel_error(e, io))
}

/// Whether a gfx950 device answers (the routing condition).
pub fn gpu_available() -> bool {
    gzgpu_sys::thread_ctx().is_ok()
}
