/*
 * gzoracle — CPU restatement of gzippy's parallel single-member gzip decode
 * path, used ONLY as the checker for the HIP implementation (tests, smoke and
 * the bench's cpu_baseline leg).  Nothing in the product links or calls it.
 *
 * Each function cites the reference file:line whose behaviour it restates
 * (paths relative to the gzippy source tree).
 */
#ifndef GZORACLE_H
#define GZORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (mirror ParallelError, single_member.rs:1384-1413) */
enum {
    GZO_OK = 0,
    GZO_INVALID_HEADER = 1,
    GZO_INVALID_FORMAT = 2,
    GZO_DECODE_FAILED = 3,
    GZO_SIZE_MISMATCH = 4,
    GZO_CRC_MISMATCH = 5,
    GZO_OUTPUT_FULL = 6,
};

/* ---- CRC32 (crc32.rs:44-160) ---- */
uint32_t gzo_crc32(uint32_t crc, const uint8_t *p, size_t n);
uint32_t gzo_mulmod(uint32_t a, uint32_t b);
uint32_t gzo_x_pow_mod(uint64_t exponent);
uint32_t gzo_crc32_combine(uint32_t crc1, uint32_t crc2, uint64_t len2);

/* ---- RFC 1952 envelope (gzip_format.rs:47-160) ---- */
int gzo_parse_header(const uint8_t *p, size_t n, size_t *hdr_len);

/* ---- inflate with block-boundary recording ---- */
typedef struct {
    uint64_t bit;      /* bit offset of the block header */
    uint64_t out_off;  /* output offset (elements) where the block's data begins */
    uint32_t btype;
    uint32_t bfinal;
} gzo_block;

typedef struct {
    uint64_t end_bit;   /* bit just past the last decoded block (or the stop block's header) */
    uint64_t out_len;   /* elements written */
    int32_t status;     /* GZO_OK / GZO_DECODE_FAILED / GZO_OUTPUT_FULL */
    int32_t bfinal_hit; /* 1 when the final block was decoded */
    uint64_t n_blocks;  /* blocks seen (may exceed the caller's capacity) */
} gzo_result;

/*
 * Decode raw DEFLATE from `start_bit` until either the final block ends, or a
 * non-final block whose header bit is >= stop_bit is reached (that block is
 * not decoded; end_bit = its header bit).  Stop rule: chunk_decode.rs:2029-2040.
 *
 * marker_mode == 0: back-references resolve against `window` (32768 bytes,
 *   oldest first; NULL = all zeros, like WindowMap{0: zeros},
 *   chunk_fetcher.rs:694-696); output goes to out8.
 * marker_mode == 1: window unknown; output is u16 in out16, literals < 256 and
 *   markers 32768 + window index (replace_markers.rs:6-23,
 *   marker_inflate.rs:255-268).
 */
int gzo_inflate(const uint8_t *in, size_t in_len, uint64_t start_bit, uint64_t stop_bit,
                int marker_mode, const uint8_t *window, uint8_t *out8, uint16_t *out16,
                size_t out_cap, gzo_block *blocks, size_t blocks_cap, gzo_result *res);

/* MapMarkers: out[i] = v < 256 ? v : window[v - 32768]  (replace_markers.rs:47,129) */
int gzo_resolve_markers(const uint16_t *in, size_t n, const uint8_t *window, uint8_t *out);

/* ---- block finder (blockfinder_validation.rs) ---- */
typedef struct {
    uint64_t bit_offset; /* earliest valid start */
    uint64_t seek_bit;   /* where decoding starts */
    uint32_t kind;       /* 0 stored, 2 dynamic */
    uint8_t hlit, hdist, hclen, pad;
} gzo_boundary;

/* dynamic-Huffman candidate predicate at one bit offset (scalar scan :631-749) */
int gzo_is_dynamic_candidate(const uint8_t *in, size_t in_len, uint64_t bit);
/* find_blocks(start_bit, end_bit), sorted by (bit_offset, seek_bit) (:520-527) */
size_t gzo_find_blocks(const uint8_t *in, size_t in_len, uint64_t start_bit, uint64_t end_bit,
                       gzo_boundary *out, size_t cap);

/* ---- whole-stream serial decode (drive_thin_t1_oracle, chunk_fetcher.rs:512;
 *      multi-member walk, mod.rs:725-741) ---- */
int gzo_gunzip(const uint8_t *in, size_t in_len, uint8_t *out, size_t out_cap, size_t *out_len,
               size_t *members);

/* ---- threaded speculative decode of one member (gzpar.c; chunk_fetcher.rs
 *      :638-1600 ParallelSM with `threads` workers and `chunk_bytes` partitions) ---- */
int gzo_gunzip_parallel(const uint8_t *in, size_t in_len, uint8_t *out, size_t out_cap, size_t *out_len,
                        int threads, size_t chunk_bytes);

#ifdef __cplusplus
}
#endif
#endif
