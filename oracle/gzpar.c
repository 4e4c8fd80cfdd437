/*
 * gzpar.c — TEST INFRASTRUCTURE.  A threaded CPU restatement of gzippy's
 * speculative parallel single-member decode (ParallelSM), built from the
 * oracle's pieces, used only as the bench's multi-core CPU baseline and as a
 * second checker in the tests.  The product never links, loads or calls it.
 *
 * Follows chunk_fetcher.rs (src/decompress/parallel/):
 *   - partition grid of chunk_bytes compressed bytes (GzipBlockFinder,
 *     gzip_block_finder.rs:54; the chunk size of adjusted_chunk_size_bytes,
 *     single_member.rs:128);
 *   - per partition, the block finder over at most 512 KiB after the
 *     partition start and a speculative window-absent (marker) decode from
 *     the first candidate that decodes cleanly up to the next partition's
 *     start (try_speculative_decode_candidate :2506,
 *     speculative_decode_find_boundary :2616, stop rule chunk_decode.rs:2029);
 *     partition 0 decodes from the deflate start with the zero window
 *     (WindowMap{0: zeros}, :694-696);
 *   - in order: a chunk is accepted when it starts where its predecessor
 *     ended, else it is decoded again from that end (:2506-2729);
 *   - end windows chained chunk to chunk (get_last_window,
 *     chunk_data.rs:1295), markers resolved in parallel
 *     (replace_markers.rs:129);
 *   - CRC32 per chunk in parallel, combined (crc32.rs:127), checked with the
 *     size against the trailer (sm_driver.rs:212-218).
 * T worker threads (pthreads) pull chunks from an atomic counter, like the
 * reference's worker pool (chunk_fetcher.rs:768-802).
 */
#include "gzoracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define WIN 32768u
#define SCAN_BYTES (512u * 1024u)

typedef struct {
    uint64_t start_bit, end_bit, out_len, out_pos;
    int marker;        /* u16 entries (markers) in buf16, else bytes in buf8 */
    int bfinal;
    int status;
    uint16_t *buf16;
    uint8_t *buf8;
    size_t cap;
    uint32_t crc;
} chunk_t;

typedef struct {
    const uint8_t *in;
    size_t n;
    uint64_t ds, de;
    uint64_t *grid;   /* partition starts (bits), grid[k] for k = 0..nchunk */
    chunk_t *ch;
    size_t nchunk;
    atomic_size_t next;
    uint8_t *out;
    size_t out_cap;
    double ratio;     /* expected output bytes per compressed byte */
    int phase;        /* 0 speculative decode, 1 resolve + CRC */
} job_t;

/* decode chunk k from `start` (marker mode unless k == 0), growing the
 * buffer until it fits */
static int decode_from(job_t *J, size_t k, uint64_t start) {
    chunk_t *c = &J->ch[k];
    const uint64_t stop = k + 1 < J->nchunk ? J->grid[k + 1] : UINT64_MAX;
    const int marker = k != 0;
    const uint64_t to = stop == UINT64_MAX ? J->de : stop;
    size_t cap = c->cap ? c->cap : (size_t)((to > start ? to - start : 0) / 8 * J->ratio * 1.3) + 65536;
    for (;;) {
        if (cap > c->cap) {
            free(c->buf16);
            free(c->buf8);
            c->buf16 = NULL;
            c->buf8 = NULL;
            if (marker) c->buf16 = (uint16_t *)malloc(cap * 2);
            else c->buf8 = (uint8_t *)malloc(cap);
            if (!c->buf16 && !c->buf8) return GZO_OUTPUT_FULL;
            c->cap = cap;
        }
        gzo_result r;
        int st = gzo_inflate(J->in, J->n, start, stop, marker, NULL, c->buf8, c->buf16, c->cap, NULL, 0, &r);
        if (st == GZO_OUTPUT_FULL) {
            cap = c->cap * 2;
            continue;
        }
        c->start_bit = start;
        c->end_bit = r.end_bit;
        c->out_len = r.out_len;
        c->marker = marker;
        c->bfinal = r.bfinal_hit;
        c->status = st;
        return st;
    }
}

/* partition k: the first block-finder candidate at or after its grid bit
 * that decodes cleanly to the stop rule */
static void speculate(job_t *J, size_t k) {
    chunk_t *c = &J->ch[k];
    if (k == 0) {
        decode_from(J, 0, J->ds);
        return;
    }
    const uint64_t g = J->grid[k];
    uint64_t end = g + 8ull * SCAN_BYTES;
    if (end > J->de) end = J->de;
    gzo_boundary cand[256];
    c->status = GZO_DECODE_FAILED;
    c->start_bit = UINT64_MAX;
    /* candidates window by window (64 Kibit, as with_sync_boundary_search,
     * async_block_finder.rs:171): the first one that decodes wins */
    for (uint64_t w = g; w < end; w += 65536) {
        const uint64_t we = w + 65536 < end ? w + 65536 : end;
        size_t nc = gzo_find_blocks(J->in, J->n, w, we, cand, 256);
        if (nc > 256) nc = 256;
        for (size_t i = 0; i < nc; i++)
            if (decode_from(J, k, cand[i].seek_bit) == GZO_OK) return;
    }
    c->status = GZO_DECODE_FAILED;
    c->start_bit = UINT64_MAX;
}

static void resolve_chunk(job_t *J, size_t k) {
    chunk_t *c = &J->ch[k];
    uint8_t *dst = J->out + c->out_pos;
    if (c->marker) {
        uint8_t win[WIN];
        /* the window before the chunk: the last 32 KiB of output before it
         * (zeros before the stream start), final once the chain ran */
        const uint64_t have = c->out_pos < WIN ? c->out_pos : WIN;
        memset(win, 0, WIN - have);
        memcpy(win + (WIN - have), J->out + c->out_pos - have, have);
        /* the chunk's own last 32 KiB were resolved by the chain */
        const uint64_t body = c->out_len > WIN ? c->out_len - WIN : 0;
        gzo_resolve_markers(c->buf16, body, win, dst);
    }
    c->crc = gzo_crc32(0, dst, c->out_len);
}

static void *worker(void *arg) {
    job_t *J = (job_t *)arg;
    for (;;) {
        size_t k = atomic_fetch_add(&J->next, 1);
        if (k >= J->nchunk) break;
        if (J->phase == 0) speculate(J, k);
        else if (J->ch[k].status == GZO_OK) resolve_chunk(J, k);
    }
    return NULL;
}

static void run_pool(job_t *J, int threads) {
    atomic_store(&J->next, 0);
    pthread_t tid[256];
    int nt = threads < 1 ? 1 : threads > 256 ? 256 : threads;
    for (int t = 1; t < nt; t++) pthread_create(&tid[t], NULL, worker, J);
    worker(J);
    for (int t = 1; t < nt; t++) pthread_join(tid[t], NULL);
}

int gzo_gunzip_parallel(const uint8_t *in, size_t n, uint8_t *out, size_t out_cap, size_t *out_len, int threads,
                        size_t chunk_bytes) {
    *out_len = 0;
    if (n < 2 || in[0] != 0x1f || in[1] != 0x8b) return GZO_OK;
    size_t hl;
    int st = gzo_parse_header(in, n, &hl);
    if (st) return st;
    if (n < hl + 8) return GZO_INVALID_FORMAT;
    const uint8_t *t = in + n - 8;
    const uint32_t exp_crc = (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16 | (uint32_t)t[3] << 24;
    const uint32_t exp_isize = (uint32_t)t[4] | (uint32_t)t[5] << 8 | (uint32_t)t[6] << 16 | (uint32_t)t[7] << 24;
    job_t J;
    memset(&J, 0, sizeof J);
    J.in = in;
    J.n = n - 8;  /* the deflate stream ends before the trailer */
    J.ds = (uint64_t)hl * 8;
    J.de = (uint64_t)(n - 8) * 8;
    J.out = out;
    J.out_cap = out_cap;
    J.ratio = (double)exp_isize / (double)(n - hl - 8 + 1);
    if (chunk_bytes < 32768) chunk_bytes = 32768;
    const uint64_t cb = 8ull * chunk_bytes;
    J.nchunk = (size_t)((J.de - J.ds + cb - 1) / cb);
    if (J.nchunk == 0) J.nchunk = 1;
    J.grid = (uint64_t *)calloc(J.nchunk + 1, sizeof(uint64_t));
    J.ch = (chunk_t *)calloc(J.nchunk, sizeof(chunk_t));
    for (size_t k = 0; k <= J.nchunk; k++) J.grid[k] = J.ds + k * cb;
    J.phase = 0;
    run_pool(&J, threads);

    /* in order: accept or re-decode, output positions, end windows */
    int rc = GZO_OK;
    uint64_t pos = 0, cur = J.ds;
    size_t used = 0;
    for (size_t k = 0; k < J.nchunk; k++) {
        chunk_t *c = &J.ch[k];
        if (c->start_bit != cur || c->status != GZO_OK) {
            if (k == 0 || decode_from(&J, k, cur) != GZO_OK) {
                rc = GZO_DECODE_FAILED;
                break;
            }
        }
        c->out_pos = pos;
        if (pos + c->out_len > out_cap) {
            rc = GZO_OUTPUT_FULL;
            break;
        }
        /* end window: the chunk's last 32 KiB (or all of it) resolved now,
         * so the next chunk's markers read final bytes */
        if (c->marker) {
            uint8_t win[WIN];
            const uint64_t have = pos < WIN ? pos : WIN;
            memset(win, 0, WIN - have);
            memcpy(win + (WIN - have), out + pos - have, have);
            const uint64_t body = c->out_len > WIN ? c->out_len - WIN : 0;
            gzo_resolve_markers(c->buf16 + body, c->out_len - body, win, out + pos + body);
        } else {
            memcpy(out + pos, c->buf8, c->out_len);
        }
        pos += c->out_len;
        cur = c->end_bit;
        used = k + 1;
        if (c->bfinal) break;
    }
    if (rc == GZO_OK && (used == 0 || !J.ch[used - 1].bfinal)) rc = GZO_DECODE_FAILED;
    if (rc == GZO_OK) {
        J.nchunk = used;
        J.phase = 1;
        run_pool(&J, threads);
        uint32_t crc = 0;
        for (size_t k = 0; k < used; k++) crc = gzo_crc32_combine(crc, J.ch[k].crc, J.ch[k].out_len);
        *out_len = pos;
        if ((uint32_t)pos != exp_isize) rc = GZO_SIZE_MISMATCH;
        else if (crc != exp_crc) rc = GZO_CRC_MISMATCH;
    }
    for (size_t k = 0; k < (J.nchunk > used ? J.nchunk : used); k++) {
        free(J.ch[k].buf16);
        free(J.ch[k].buf8);
    }
    free(J.ch);
    free(J.grid);
    return rc;
}
