//! Raw binding of the C ABI in `include/gzgpu.h` (libgzgpu, gfx950), plus a
//! thin safe layer: a per-thread context and the two host-buffer calls that
//! gzippy's `decompress` / `decompress_to_writer` need.  Not compiled in this
//! repository (no Rust toolchain in its image); the declarations follow the
//! header field for field.
#![allow(non_camel_case_types)]

use std::cell::Cell;
use std::ffi::{c_char, c_int, c_void, CStr};

pub const GZGPU_OK: c_int = 0;
pub const GZGPU_INVALID_HEADER: c_int = 1;
pub const GZGPU_INVALID_FORMAT: c_int = 2;
pub const GZGPU_DECODE_FAILED: c_int = 3;
pub const GZGPU_SIZE_MISMATCH: c_int = 4;
pub const GZGPU_CRC_MISMATCH: c_int = 5;
pub const GZGPU_IO: c_int = 6;
pub const GZGPU_GPU_UNAVAILABLE: c_int = 7;
pub const GZGPU_OUT_OF_DEVICE_MEMORY: c_int = 8;
pub const GZGPU_OUTPUT_TOO_SMALL: c_int = 9;
pub const GZGPU_INVALID_ARGUMENT: c_int = 10;
pub const GZGPU_PENDING: c_int = 11;

#[repr(C)]
pub struct gzgpu_ctx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct gzgpu_server {
    _p: [u8; 0],
}
#[repr(C)]
pub struct gzgpu_batch_graph {
    _p: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct gzgpu_tuning {
    pub chunk_bytes: u64,
    pub pieces_per_chunk: u32,
    pub waves_per_cu: u32,
    pub batch_split_bytes: u64,
    pub miss_every: u32,
    pub lut_bits: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct gzgpu_stats {
    pub ms_total: f64,
    pub ms_find: f64,
    pub ms_inflate: f64,
    pub ms_resolve: f64,
    pub ms_crc: f64,
    pub compressed_bytes: u64,
    pub decompressed_bytes: u64,
    pub pieces: u32,
    pub redecoded: u32,
    pub candidates: u32,
    pub launches: u32,
    pub ms_inflate_kernel: f64,
    pub inflate_launches: u32,
    pub pad_: u32,
    pub ms_inflate_marker: f64,
    pub ms_inflate_clean: f64,
    pub clean_in_bytes: u64,
    pub clean_out_bytes: u64,
    pub marker_in_bytes: u64,
    pub marker_out_bytes: u64,
    pub midblock_pieces: u32,
    pub midblock_linked: u32,
    pub stored_members: u32,
    pub stored_blocks: u32,
    pub final_block_pieces: u32,
    pub stored_runs_skipped: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct gzgpu_boundary {
    pub bit_offset: u64,
    pub seek_bit: u64,
    pub kind: u32,
    pub info: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct gzgpu_piece_result {
    pub end_bit: u64,
    pub out_len: u64,
    pub status: i32,
    pub bfinal_hit: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct gzgpu_server_config {
    pub hbm_budget_bytes: u64,
    pub max_batch_requests: u32,
    pub workers: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct gzgpu_request_info {
    pub status: i32,
    pub members: u32,
    pub out_len: u64,
    pub queue_ms: f64,
    pub latency_ms: f64,
    pub batch: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct gzgpu_server_stats {
    pub requests_done: u64,
    pub batches: u64,
    pub bytes_in: u64,
    pub bytes_out: u64,
    pub fallbacks: u64,
    pub hbm_peak_bytes: u64,
}

pub type gzgpu_write_fn = extern "C" fn(user: *mut c_void, data: *const u8, len: usize) -> c_int;

extern "C" {
    pub fn gzgpu_ctx_create(device: c_int, out: *mut *mut gzgpu_ctx) -> c_int;
    pub fn gzgpu_ctx_destroy(ctx: *mut gzgpu_ctx);
    pub fn gzgpu_set_tuning(ctx: *mut gzgpu_ctx, t: *const gzgpu_tuning) -> c_int;
    pub fn gzgpu_get_tuning(ctx: *mut gzgpu_ctx, t: *mut gzgpu_tuning) -> c_int;
    pub fn gzgpu_last_error(ctx: *mut gzgpu_ctx) -> *const c_char;
    pub fn gzgpu_status_string(status: c_int) -> *const c_char;

    pub fn gzgpu_decompress(ctx: *mut gzgpu_ctx, input: *const u8, in_len: usize, out: *mut u8, out_cap: usize,
                            out_len: *mut usize, stats: *mut gzgpu_stats) -> c_int;
    pub fn gzgpu_decompress_to_writer(ctx: *mut gzgpu_ctx, input: *const u8, in_len: usize, write: gzgpu_write_fn,
                                      user: *mut c_void, total: *mut u64, stats: *mut gzgpu_stats) -> c_int;
    pub fn gzgpu_decompress_device(ctx: *mut gzgpu_ctx, d_in: *const c_void, in_len: usize, d_out: *mut c_void,
                                   out_cap: usize, out_len: *mut usize, crc_out: *mut u32,
                                   stats: *mut gzgpu_stats) -> c_int;
    pub fn gzgpu_decompress_device_noalloc(ctx: *mut gzgpu_ctx, d_in: *const c_void, in_len: usize,
                                           d_out: *mut c_void, out_cap: usize, out_len: *mut usize,
                                           crc_out: *mut u32, stats: *mut gzgpu_stats) -> c_int;
    pub fn gzgpu_decompress_batch(ctx: *mut gzgpu_ctx, d_in: *const c_void, in_off: *const u64, in_len: *const u64,
                                  d_out: *mut c_void, out_off: *const u64, out_cap: *const u64, n: usize,
                                  status: *mut i32, out_len: *mut u64, crc: *mut u32,
                                  stats: *mut gzgpu_stats) -> c_int;
    pub fn gzgpu_autotune(ctx: *mut gzgpu_ctx, d_in: *const c_void, in_off: *const u64, in_len: *const u64,
                          d_out: *mut c_void, out_off: *const u64, out_cap: *const u64, n: usize, budget_ms: u32,
                          best: *mut gzgpu_tuning, best_ms: *mut f64) -> c_int;
    pub fn gzgpu_batch_graph_create(ctx: *mut gzgpu_ctx, d_in: *const c_void, in_off: *const u64,
                                    in_len: *const u64, d_out: *mut c_void, out_off: *const u64,
                                    out_cap: *const u64, n: usize, out: *mut *mut gzgpu_batch_graph) -> c_int;
    pub fn gzgpu_batch_graph_launch(g: *mut gzgpu_batch_graph, status: *mut i32, out_len: *mut u64,
                                    crc: *mut u32, stats: *mut gzgpu_stats) -> c_int;
    pub fn gzgpu_batch_graph_destroy(g: *mut gzgpu_batch_graph);
    pub fn gzgpu_find_blocks(ctx: *mut gzgpu_ctx, d_in: *const c_void, in_len: usize, start_bit: u64,
                             end_bit: u64, out: *mut gzgpu_boundary, cap: usize, count: *mut usize) -> c_int;
    pub fn gzgpu_crc32(ctx: *mut gzgpu_ctx, d_data: *const c_void, len: usize, crc: *mut u32) -> c_int;
    pub fn gzgpu_inflate_piece(ctx: *mut gzgpu_ctx, d_in: *const c_void, in_len: usize, start_bit: u64,
                               stop_bit: u64, window: *const u8, d_out: *mut c_void, out_cap_elems: usize,
                               res: *mut gzgpu_piece_result) -> c_int;

    pub fn gzgpu_server_create(device: c_int, cfg: *const gzgpu_server_config, out: *mut *mut gzgpu_server)
                               -> c_int;
    pub fn gzgpu_server_destroy(s: *mut gzgpu_server);
    pub fn gzgpu_submit(s: *mut gzgpu_server, input: *const u8, in_len: usize, ticket: *mut u64) -> c_int;
    pub fn gzgpu_poll(s: *mut gzgpu_server, ticket: u64, timeout_ms: i32, info: *mut gzgpu_request_info) -> c_int;
    pub fn gzgpu_fetch(s: *mut gzgpu_server, ticket: u64, out: *mut u8, cap: usize, out_len: *mut usize) -> c_int;
    pub fn gzgpu_server_get_stats(s: *mut gzgpu_server, st: *mut gzgpu_server_stats) -> c_int;
}

// ---------------------------------------------------------------- safe layer

/// A failed call: the status code and the context's message.
#[derive(Debug)]
pub struct GpuError {
    pub status: c_int,
    pub message: String,
}

impl std::fmt::Display for GpuError {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        let s = unsafe { CStr::from_ptr(gzgpu_status_string(self.status)) }.to_string_lossy();
        write!(f, "{s}: {}", self.message)
    }
}

impl std::error::Error for GpuError {}

thread_local! {
    // one context per calling thread: the header's threading contract
    static CTX: Cell<*mut gzgpu_ctx> = const { Cell::new(std::ptr::null_mut()) };
}

fn error_of(ctx: *mut gzgpu_ctx, status: c_int) -> GpuError {
    let message = unsafe { CStr::from_ptr(gzgpu_last_error(ctx)) }.to_string_lossy().into_owned();
    GpuError { status, message }
}

/// The calling thread's context on device 0 (created on first use; lives
/// until the thread ends).
pub fn thread_ctx() -> Result<*mut gzgpu_ctx, GpuError> {
    CTX.with(|c| {
        if c.get().is_null() {
            let mut p = std::ptr::null_mut();
            let st = unsafe { gzgpu_ctx_create(0, &mut p) };
            if st != GZGPU_OK {
                return Err(GpuError { status: st, message: "no gfx950 device".into() });
            }
            c.set(p);
        }
        Ok(c.get())
    })
}

struct Sink<'a> {
    w: &'a mut dyn std::io::Write,
    err: Option<std::io::Error>,
}

extern "C" fn write_cb(user: *mut c_void, data: *const u8, len: usize) -> c_int {
    let s = unsafe { &mut *(user as *mut Sink) };
    match s.w.write_all(unsafe { std::slice::from_raw_parts(data, len) }) {
        Ok(()) => 0,
        Err(e) => {
            s.err = Some(e);
            1
        }
    }
}

/// Decode a gzip stream into `writer` in order (partial output may precede a
/// CRC or size error).  Ok(bytes written).  A writer error comes back as
/// GZGPU_IO with the io::Error in `io`.
pub fn decompress_to_writer(data: &[u8], writer: &mut dyn std::io::Write)
                            -> Result<u64, (GpuError, Option<std::io::Error>)> {
    let ctx = thread_ctx().map_err(|e| (e, None))?;
    let mut sink = Sink { w: writer, err: None };
    let mut total = 0u64;
    let st = unsafe {
        gzgpu_decompress_to_writer(ctx, data.as_ptr(), data.len(), write_cb, &mut sink as *mut Sink as *mut c_void,
                                   &mut total, std::ptr::null_mut())
    };
    if st == GZGPU_OK {
        Ok(total)
    } else {
        Err((error_of(ctx, st), sink.err.take()))
    }
}

/// Decode a gzip stream into a new Vec.
pub fn decompress(data: &[u8]) -> Result<Vec<u8>, GpuError> {
    let ctx = thread_ctx()?;
    let mut cap = if data.len() >= 18 {
        u32::from_le_bytes([data[data.len() - 4], data[data.len() - 3], data[data.len() - 2], data[data.len() - 1]])
            as usize
            + 1
    } else {
        1 << 16
    };
    loop {
        let mut out = vec![0u8; cap];
        let mut n = 0usize;
        let st = unsafe {
            gzgpu_decompress(ctx, data.as_ptr(), data.len(), out.as_mut_ptr(), cap, &mut n, std::ptr::null_mut())
        };
        if st == GZGPU_OUTPUT_TOO_SMALL && n > cap {
            cap = n;
            continue;
        }
        if st != GZGPU_OK {
            return Err(error_of(ctx, st));
        }
        out.truncate(n);
        return Ok(out);
    }
}
