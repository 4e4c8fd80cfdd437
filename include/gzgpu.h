/*
 * gzgpu.h — C ABI of the MI355X (gfx950) parallel gzip decompressor.
 *
 * The library replaces gzippy's parallel single-member decode path behind
 * gzippy::decompress / decompress_to_writer.  Every entry point takes plain
 * pointers and sizes; none retains a caller buffer past the call.
 *
 * Reference interfaces replaced (paths into the gzippy tree):
 *   gzgpu_decompress            src/lib.rs:176 decompress / :189 decompress_with_threads
 *                               → src/decompress/mod.rs:317 decompress_bytes
 *   gzgpu_decompress_to_writer  src/lib.rs:204 decompress_to_writer / :221 ..._with_threads
 *   gzgpu_decompress_device     src/decompress/parallel/single_member.rs:987
 *                               decompress_parallel (the ParallelSM arm,
 *                               src/decompress/mod.rs:629-650), device-resident
 *   gzgpu_decompress_batch      src/decompress/bgzf.rs:2370
 *                               decompress_multi_member_parallel_to_vec (member-per-worker)
 *   gzgpu_find_blocks           src/decompress/parallel/blockfinder_validation.rs:483 find_blocks
 *   gzgpu_crc32                 src/decompress/parallel/crc32.rs:155 crc32_fold
 *   gzgpu_status_string         src/decompress/parallel/single_member.rs:1415-1437 Display
 *                               for ParallelError
 */
#ifndef GZGPU_H
#define GZGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1..6 map 1:1 onto ParallelError
 * (src/decompress/parallel/single_member.rs:1384-1413); the rest are the
 * device-side conditions a CPU build cannot hit. */
typedef enum {
    GZGPU_OK = 0,
    GZGPU_INVALID_HEADER = 1,      /* ParallelError::InvalidHeader */
    GZGPU_INVALID_FORMAT = 2,      /* ParallelError::InvalidGzipFormat */
    GZGPU_DECODE_FAILED = 3,       /* ParallelError::DecodeFailed */
    GZGPU_SIZE_MISMATCH = 4,       /* ParallelError::SizeMismatch */
    GZGPU_CRC_MISMATCH = 5,        /* ParallelError::CrcMismatch */
    GZGPU_IO = 6,                  /* ParallelError::Io (writer callback failed) */
    GZGPU_GPU_UNAVAILABLE = 7,     /* no gfx950 device / HIP runtime error */
    GZGPU_OUT_OF_DEVICE_MEMORY = 8,
    GZGPU_OUTPUT_TOO_SMALL = 9,    /* out_cap < decoded size; *out_len holds the size needed */
    GZGPU_INVALID_ARGUMENT = 10,
    GZGPU_PENDING = 11,            /* gzgpu_poll: the request is not finished yet */
} gzgpu_status;

typedef struct gzgpu_ctx gzgpu_ctx;

/* Tunables; a 0 field is automatic and gzgpu_get_tuning returns it as 0.
 * pieces_per_chunk 0: pieces sized so that each resident wave slot gets an
 * equal share of the call's decode cost (chunk_bytes unused); otherwise a
 * fixed grid of chunk_bytes / pieces_per_chunk. */
typedef struct {
    uint64_t chunk_bytes;       /* speculation grid, compressed bytes per chunk (0: 512 KiB) */
    uint32_t pieces_per_chunk;  /* wavefronts per chunk: sub-splits at found block boundaries */
    uint32_t waves_per_cu;      /* persistent-grid occupancy cap (0: as many as LDS allows) */
    uint64_t batch_split_bytes; /* batch: split only members above this compressed size (0: by expected output) */
    uint32_t miss_every;        /* test hook: plan every n-th speculative start one bit late, so that its
                                   predecessor's end misses it and the re-decode path runs (0: off) */
    uint32_t lut_bits;          /* inflate kernel's litlen LUT root: 0 (automatic: 9), 9, or 10 (a 1.8 KB larger
                                   LUT per wavefront: 13 resident per CU instead of 16) */
} gzgpu_tuning;

/* Per-call timing and shape, filled when a non-NULL pointer is passed. */
typedef struct {
    double ms_total;            /* device time of the whole call (HIP events) */
    double ms_find;             /* block finder */
    double ms_inflate;          /* speculative inflate (all launches) */
    double ms_resolve;          /* window chain + marker resolution */
    double ms_crc;              /* CRC32 fold + combine */
    uint64_t compressed_bytes;
    uint64_t decompressed_bytes;
    uint32_t pieces;            /* speculative pieces decoded */
    uint32_t redecoded;         /* pieces re-decoded after a speculation miss or slab overflow */
    uint32_t candidates;        /* block-finder hits */
    uint32_t launches;          /* kernel launches */
    double ms_inflate_kernel;   /* summed k_inflate durations (HIP events around each launch) */
    uint32_t inflate_launches;  /* k_inflate launches (2 per speculation round: marker, clean) */
    uint32_t pad_;
    double ms_inflate_marker;   /* u16-ring kernel launches (calls with speculative pieces) */
    double ms_inflate_clean;    /* u8-ring kernel launches (calls without) */
    uint64_t clean_in_bytes;    /* compressed bytes consumed by the u8-ring kernel */
    uint64_t clean_out_bytes;   /* bytes written by the u8-ring kernel */
    uint64_t marker_in_bytes;   /* compressed bytes consumed by the u16-ring kernel */
    uint64_t marker_out_bytes;  /* bytes written by the u16-ring kernel (u16 entries count 2) */
    uint32_t midblock_pieces;   /* pieces planned to start inside a block (at a sync point) */
    uint32_t midblock_linked;   /* of those, linked to their predecessor's exact stop */
    uint32_t stored_members;    /* members copied by the stored-block split (no speculation) */
    uint32_t stored_blocks;     /* their stored blocks */
    uint32_t final_block_pieces; /* pieces planned to start inside a member's final block */
    uint32_t stored_runs_skipped; /* stored blocks inside split members whose payload the block finder skipped */
} gzgpu_stats;

/* ---- context ---------------------------------------------------------- */
/* One context per calling thread (or lock around it).  Holds the HIP stream,
 * grown device buffers and tuning. */
int gzgpu_ctx_create(int device, gzgpu_ctx **out);
void gzgpu_ctx_destroy(gzgpu_ctx *ctx);
int gzgpu_set_tuning(gzgpu_ctx *ctx, const gzgpu_tuning *t);
int gzgpu_get_tuning(gzgpu_ctx *ctx, gzgpu_tuning *t);
const char *gzgpu_last_error(gzgpu_ctx *ctx);
const char *gzgpu_status_string(int status);

/* ---- whole-stream decode (host buffers) -------------------------------- */
/* Non-gzip input (no 1f 8b magic) returns GZGPU_OK with *out_len = 0
 * (src/lib.rs:50-52, src/decompress/mod.rs:337-339).  Multi-member input is
 * the concatenation of member outputs; bytes after a valid member that do not
 * start a new member end the walk cleanly (src/decompress/mod.rs:736-741). */
int gzgpu_decompress(gzgpu_ctx *ctx, const uint8_t *in, size_t in_len, uint8_t *out, size_t out_cap,
                     size_t *out_len, gzgpu_stats *stats);

/* Streaming sink: `write` is called in order with consecutive output pieces;
 * a non-zero return aborts with GZGPU_IO.  Partial output may have been
 * delivered before a CRC/ISIZE error (single_member.rs:17-22). */
typedef int (*gzgpu_write_fn)(void *user, const uint8_t *data, size_t len);
int gzgpu_decompress_to_writer(gzgpu_ctx *ctx, const uint8_t *in, size_t in_len, gzgpu_write_fn write,
                               void *user, uint64_t *total, gzgpu_stats *stats);

/* ---- device-resident single member -------------------------------------- */
/* `d_in` and `d_out` are device pointers (d_in readable for in_len bytes).
 * Decodes one gzip member into d_out; verifies ISIZE and CRC32 against the
 * trailer.  `*out_len` receives the decoded size (also on
 * GZGPU_OUTPUT_TOO_SMALL). */
int gzgpu_decompress_device(gzgpu_ctx *ctx, const void *d_in, size_t in_len, void *d_out, size_t out_cap,
                            size_t *out_len, uint32_t *crc_out, gzgpu_stats *stats);

/* ---- device-resident single member, no allocation ------------------------ */
/* gzgpu_decompress_device without any allocation inside the call: no
 * hipMalloc, hipFree, hipHostMalloc or other allocating routine runs.  Every
 * device buffer the decode needs comes from one static array per device
 * (256 MiB, held from library load to process exit), host staging from a
 * pinned block made by gzgpu_ctx_create.  For callers that cannot allocate:
 * serving loops with a fixed memory footprint, harnesses that forbid
 * allocation.  Limits: in_len <= GZGPU_NOALLOC_MAX_IN_BYTES and the member's
 * trailer ISIZE <= GZGPU_NOALLOC_MAX_OUT_BYTES (GZGPU_INVALID_ARGUMENT
 * otherwise, before any device work); a member inside them whose decode still
 * needs more scratch fails with GZGPU_OUT_OF_DEVICE_MEMORY.  Calls that
 * overlap on one device (two contexts, two threads) share the array and race:
 * serialise them.  Output equals gzgpu_decompress_device's.  No counterpart
 * in the reference. */
#define GZGPU_NOALLOC_MAX_IN_BYTES (32ull << 20)
#define GZGPU_NOALLOC_MAX_OUT_BYTES (32ull << 20)
int gzgpu_decompress_device_noalloc(gzgpu_ctx *ctx, const void *d_in, size_t in_len, void *d_out, size_t out_cap,
                                    size_t *out_len, uint32_t *crc_out, gzgpu_stats *stats);

/* ---- batched members (one persistent launch) ---------------------------- */
/* n independent gzip members, member i at d_in + in_off[i] (in_len[i]
 * bytes); output of member i goes to d_out + out_off[i] with room for
 * out_cap[i] bytes.  status[i], out_len[i] and crc[i] are filled per
 * member; the return value is GZGPU_OK when every member decoded and
 * verified, else the first failing member's status. */
int gzgpu_decompress_batch(gzgpu_ctx *ctx, const void *d_in, const uint64_t *in_off, const uint64_t *in_len,
                           void *d_out, const uint64_t *out_off, const uint64_t *out_cap, size_t n,
                           int32_t *status, uint64_t *out_len, uint32_t *crc, gzgpu_stats *stats);

/* ---- autotuner ---------------------------------------------------------- */
/* Times candidate tunings (piece grid, wavefronts per chunk, waves per CU,
 * batch split rule) on the caller's own members, laid out as for
 * gzgpu_decompress_batch (n = 1 for a single member), keeps the fastest by
 * device time (best of two runs after a warm-up; the planner's default unless
 * another candidate is more than 2% faster) in the context and returns it in
 * *best.  d_out is overwritten.
 * Stops after about budget_ms of device time.  No counterpart in the
 * reference, whose chunk size comes from a fixed formula
 * (adjusted_chunk_size_bytes, single_member.rs:128). */
int gzgpu_autotune(gzgpu_ctx *ctx, const void *d_in, const uint64_t *in_off, const uint64_t *in_len, void *d_out,
                   const uint64_t *out_off, const uint64_t *out_cap, size_t n, uint32_t budget_ms,
                   gzgpu_tuning *best, double *best_ms);

/* ---- batch as a replayable hipGraph ------------------------------------- */
/* Captures the whole decode of a fixed batch layout (members as for
 * gzgpu_decompress_batch) into one hipGraph, so a launch is one
 * hipGraphLaunch and one small copy back.  On the device: header parse; for
 * members whose decode cost exceeds one resident wavefront's share, the
 * block finder, a piece plan (starts at block headers or inside blocks) and
 * after the inflate the link check, with in-place re-decodes of missed
 * pieces and a whole-remainder fallback; a persistent inflate kernel fed by
 * device-side work queues (longest first); the window chain, marker
 * resolution, CRC32 and the trailer checks.  The contents of d_in may change
 * between launches; offsets, lengths and capacities may not (piece counts
 * come from the trailers at creation: any contents decode exactly).  Device
 * memory: marker slabs of about 8 x the split members' output capacity.
 * Destroy a graph before its context.  No counterpart in the reference (the
 * batched form of decompress_multi_member_parallel_to_vec, bgzf.rs:2370, and
 * of the speculative pipeline, chunk_fetcher.rs:638-1600, with the device
 * keeping every decision). */
typedef struct gzgpu_batch_graph gzgpu_batch_graph;
int gzgpu_batch_graph_create(gzgpu_ctx *ctx, const void *d_in, const uint64_t *in_off, const uint64_t *in_len,
                             void *d_out, const uint64_t *out_off, const uint64_t *out_cap, size_t n,
                             gzgpu_batch_graph **out);
/* status / out_len / crc: n entries each (NULL to skip); returns the first
 * member's error or GZGPU_OK. */
int gzgpu_batch_graph_launch(gzgpu_batch_graph *g, int32_t *status, uint64_t *out_len, uint32_t *crc,
                             gzgpu_stats *stats);
void gzgpu_batch_graph_destroy(gzgpu_batch_graph *g);

/* The plain per-member form: the same members decoded with one inflate launch
 * per member (one wavefront each, whole member), queued back to back on the
 * context's stream with no host synchronisation in between; header parse,
 * CRC32 and trailer checks on the device as in the graph.  The baseline the
 * batched forms are measured against.  No counterpart in the reference. */
int gzgpu_decompress_members_plain(gzgpu_ctx *ctx, const void *d_in, const uint64_t *in_off, const uint64_t *in_len,
                                   void *d_out, const uint64_t *out_off, const uint64_t *out_cap, size_t n,
                                   int32_t *status, uint64_t *out_len, uint32_t *crc, gzgpu_stats *stats);

/* ---- exposed kernels (parity tests, tools) ------------------------------- */
typedef struct {
    uint64_t bit_offset; /* earliest valid start */
    uint64_t seek_bit;   /* where decoding starts */
    uint32_t kind;       /* 0 stored, 2 dynamic */
    uint32_t info;       /* dynamic: hlit | hdist << 8 | hclen << 16; stored: LEN << 16, bit 0 set
                            when the bits after the LEN-byte payload parse as a block header */
} gzgpu_boundary;

/* Every block-finder hit in bit range [start_bit, end_bit) of the device
 * buffer d_in[0..in_len), sorted by (bit_offset, seek_bit).  Returns the
 * hit count in *count even when it exceeds cap (then only cap are written). */
int gzgpu_find_blocks(gzgpu_ctx *ctx, const void *d_in, size_t in_len, uint64_t start_bit, uint64_t end_bit,
                      gzgpu_boundary *out, size_t cap, size_t *count);

/* gzip CRC32 of a device buffer (finalized value, like crc32fast / zlib). */
int gzgpu_crc32(gzgpu_ctx *ctx, const void *d_data, size_t len, uint32_t *crc);

/* Raw-DEFLATE decode of bit range [start_bit, ...) with the stop rule of
 * chunk_decode.rs:2029-2040 (stop before a non-final block whose header bit is
 * >= stop_bit).  window == NULL: marker mode, output u16 values
 * (< 256 literal, >= 32768 window marker) into d_out16; window != NULL (host,
 * 32768 bytes): clean decode into d_out8.  One wavefront, one piece: the
 * building block of the parallel path, exposed for parity tests. */
typedef struct {
    uint64_t end_bit;
    uint64_t out_len;
    int32_t status;      /* gzgpu_status */
    int32_t bfinal_hit;
} gzgpu_piece_result;
int gzgpu_inflate_piece(gzgpu_ctx *ctx, const void *d_in, size_t in_len, uint64_t start_bit, uint64_t stop_bit,
                        const uint8_t *window, void *d_out, size_t out_cap_elems, gzgpu_piece_result *res);

/* ---- seekable index ------------------------------------------------------ */
/* Random access into a single-member gzip stream (src/decompress/index.rs).
 * gzgpu_index_build (index.rs:54 build_index + :145 serialize_index) decodes
 * the member once and writes an index in index.rs's version-1 layout
 * ("GZIDX\x01", deflate offset u32, total size u64, point count u32, 2
 * reserved bytes; per point: bit offset from the deflate start u64, output
 * offset u64, 32 KiB window) with checkpoints at block starts at least
 * `interval` output bytes apart.  *index_len receives the size needed; with
 * cap below it the call returns GZGPU_OUTPUT_TOO_SMALL.  Multi-member input
 * and members that carry their size (BGZF) are refused (GZGPU_INVALID_ARGUMENT),
 * as index.rs:61-72 does.
 * gzgpu_index_seek (index.rs:184 load_index + :344 seek_decompress) writes the
 * output bytes [offset, offset + max_bytes) (clamped to the stream's end) to
 * out: the checkpoints around the range decode in parallel from their
 * windows; an index that does not belong to the input is an error. */
int gzgpu_index_build(gzgpu_ctx *ctx, const uint8_t *in, size_t in_len, uint64_t interval, uint8_t *index,
                      size_t cap, size_t *index_len);
int gzgpu_index_seek(gzgpu_ctx *ctx, const uint8_t *in, size_t in_len, const uint8_t *index, size_t index_len,
                     uint64_t offset, uint64_t max_bytes, uint8_t *out, size_t out_cap, size_t *out_len);

/* ---- serving: continuous batching --------------------------------------- */
/* A request is one gzip stream (one member or several concatenated, like
 * gzippy::decompress on that buffer: src/lib.rs:176, src/decompress/mod.rs:725-741).
 * gzgpu_submit copies the input and returns at once.  Worker threads (one HIP
 * stream each) keep admitting every pending request that fits the HBM budget
 * into the next device batch, so requests that arrive while others run on the
 * GPU join the following batch and batches overlap on the device.  Each
 * result is byte-identical to decoding that request alone.  No counterpart in
 * the reference, which has no serving loop; it replaces a host-side request
 * loop calling decompress() per request. */
typedef struct gzgpu_server gzgpu_server;

typedef struct {
    uint64_t hbm_budget_bytes;   /* device bytes held for admitted requests (default 200 GB) */
    uint32_t max_batch_requests; /* requests per device batch (default 4096) */
    uint32_t workers;            /* batches in flight, one stream each (default 2) */
} gzgpu_server_config;

typedef struct {
    int32_t status;              /* gzgpu_status of the request's decode */
    uint32_t members;            /* gzip members in the request */
    uint64_t out_len;            /* decoded bytes */
    double queue_ms;             /* submit -> its batch started */
    double latency_ms;           /* submit -> result ready */
    uint64_t batch;              /* sequence number of the batch that ran it */
} gzgpu_request_info;

typedef struct {
    uint64_t requests_done;
    uint64_t batches;
    uint64_t bytes_in;
    uint64_t bytes_out;
    uint64_t fallbacks;          /* requests re-decoded member by member after a split miss */
    uint64_t hbm_peak_bytes;
} gzgpu_server_stats;

int gzgpu_server_create(int device, const gzgpu_server_config *cfg, gzgpu_server **out);
void gzgpu_server_destroy(gzgpu_server *s);
int gzgpu_submit(gzgpu_server *s, const uint8_t *in, size_t in_len, uint64_t *ticket);
/* GZGPU_OK and *info when the request finished (its own status in info->status),
 * GZGPU_PENDING when it did not within timeout_ms (0: do not wait, < 0: wait). */
int gzgpu_poll(gzgpu_server *s, uint64_t ticket, int32_t timeout_ms, gzgpu_request_info *info);
/* Copy a finished request's output and release the ticket; GZGPU_OUTPUT_TOO_SMALL
 * with *out_len set (ticket kept) when cap is short. */
int gzgpu_fetch(gzgpu_server *s, uint64_t ticket, uint8_t *out, size_t cap, size_t *out_len);
int gzgpu_server_get_stats(gzgpu_server *s, gzgpu_server_stats *st);

#ifdef __cplusplus
}
#endif
#endif
