/*
 * gzoracle.c — TEST INFRASTRUCTURE.  CPU restatement of gzippy's parallel
 * gzip decode path, used only as the checker (pytest, smoke, the bench's
 * cpu_baseline).  The product never links, loads or calls this file.
 *
 * Reference citations are file:line into the gzippy tree
 * (src/decompress/parallel/ unless stated otherwise).
 */
#include "gzoracle.h"

#include <stdlib.h>
#include <string.h>

/* ===================================================================== */
/* CRC32 — reflected 0xEDB88320 (crc32.rs:44-160)                        */
/* ===================================================================== */

#define CRC_POLY 0xEDB88320u
static uint32_t crc_tab[8][256];

__attribute__((constructor)) static void crc_init(void) {
    for (uint32_t i = 0; i < 256; i++) {
        uint32_t c = i;
        for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ CRC_POLY : c >> 1;
        crc_tab[0][i] = c;
    }
    for (uint32_t i = 0; i < 256; i++)
        for (int t = 1; t < 8; t++)
            crc_tab[t][i] = (crc_tab[t - 1][i] >> 8) ^ crc_tab[0][crc_tab[t - 1][i] & 0xFF];
}

/* Finalized-CRC continuation, same contract as crc32_fold (crc32.rs:155):
 * pass 0 for an empty prefix.  Slicing-by-8. */
uint32_t gzo_crc32(uint32_t crc, const uint8_t *p, size_t n) {
    uint32_t c = ~crc;
    while (n >= 8) {
        uint32_t lo = c ^ ((uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 |
                           (uint32_t)p[3] << 24);
        uint32_t hi = (uint32_t)p[4] | (uint32_t)p[5] << 8 | (uint32_t)p[6] << 16 |
                      (uint32_t)p[7] << 24;
        c = crc_tab[7][lo & 0xFF] ^ crc_tab[6][(lo >> 8) & 0xFF] ^ crc_tab[5][(lo >> 16) & 0xFF] ^
            crc_tab[4][lo >> 24] ^ crc_tab[3][hi & 0xFF] ^ crc_tab[2][(hi >> 8) & 0xFF] ^
            crc_tab[1][(hi >> 16) & 0xFF] ^ crc_tab[0][hi >> 24];
        p += 8;
        n -= 8;
    }
    while (n--) c = (c >> 8) ^ crc_tab[0][(c ^ *p++) & 0xFF];
    return ~c;
}

/* polynomial_multiply_modulo (crc32.rs:64-80): carry-less a*b mod P, reflected. */
uint32_t gzo_mulmod(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (uint32_t m = 1u << 31; m; m >>= 1) {
        if (a & m) r ^= b;
        b = (b & 1) ? (b >> 1) ^ CRC_POLY : b >> 1;
    }
    return r;
}

/* x_power_modulo (crc32.rs:104-120): x^e mod P by square-and-multiply over
 * the X2N table (x^(2^k)), built here from x^1 = 1<<30 (reflected). */
uint32_t gzo_x_pow_mod(uint64_t e) {
    uint32_t x2n[64];
    x2n[0] = 1u << 30;
    for (int k = 1; k < 64; k++) x2n[k] = gzo_mulmod(x2n[k - 1], x2n[k - 1]);
    uint32_t p = 1u << 31; /* x^0 */
    for (int k = 0; e; k++, e >>= 1)
        if (e & 1) p = gzo_mulmod(x2n[k], p);
    return p;
}

/* combine_crc32 (crc32.rs:127-133). */
uint32_t gzo_crc32_combine(uint32_t crc1, uint32_t crc2, uint64_t len2) {
    return gzo_mulmod(gzo_x_pow_mod(len2 * 8), crc1) ^ crc2;
}

/* ===================================================================== */
/* RFC 1952 header (gzip_format.rs:47-131)                               */
/* ===================================================================== */

int gzo_parse_header(const uint8_t *p, size_t n, size_t *hdr_len) {
    if (n < 10) return GZO_INVALID_HEADER;
    if (p[0] != 0x1f || p[1] != 0x8b || p[2] != 8) return GZO_INVALID_HEADER;
    uint8_t flg = p[3];
    size_t off = 10;
    if (flg & 4) { /* FEXTRA */
        if (off + 2 > n) return GZO_INVALID_HEADER;
        size_t xlen = (size_t)p[off] | (size_t)p[off + 1] << 8;
        off += 2;
        if (off + xlen > n) return GZO_INVALID_HEADER;
        off += xlen;
    }
    for (int f = 8; f <= 16; f <<= 1) { /* FNAME then FCOMMENT, zero-terminated */
        if (flg & f) {
            while (off < n && p[off]) off++;
            if (off >= n) return GZO_INVALID_HEADER;
            off++;
        }
    }
    if (flg & 2) { /* FHCRC */
        if (off + 2 > n) return GZO_INVALID_HEADER;
        off += 2;
    }
    *hdr_len = off;
    return GZO_OK;
}

/* ===================================================================== */
/* Bit reader: LSB-first, bits past the end read as zero                 */
/* (bit_reader.rs:46-133)                                                */
/* ===================================================================== */

typedef struct {
    const uint8_t *p;
    size_t n;
    uint64_t pos; /* absolute bit position */
} bitrd;

static inline uint64_t br_peek(const bitrd *b, unsigned k) {
    /* k <= 32 */
    uint64_t byte = b->pos >> 3;
    unsigned sh = (unsigned)(b->pos & 7);
    uint64_t v = 0;
    if (byte + 8 <= b->n) {
        memcpy(&v, b->p + byte, 8);
    } else {
        for (unsigned i = 0; i < 8 && byte + i < b->n; i++) v |= (uint64_t)b->p[byte + i] << (8 * i);
    }
    v >>= sh;
    return k >= 64 ? v : (v & ((1ull << k) - 1));
}
static inline uint32_t br_read(bitrd *b, unsigned k) {
    uint32_t v = (uint32_t)br_peek(b, k);
    b->pos += k;
    return v;
}

/* ===================================================================== */
/* Canonical Huffman (RFC 1951 3.2.2) — 10-bit cache + canonical walk    */
/* ===================================================================== */

#define HROOT 10
typedef struct {
    uint16_t lut[1 << HROOT]; /* (sym << 4) | len, len == 0: not cached */
    uint16_t count[16];
    uint16_t sorted[320];
} huff_t;

/* Returns 0 on success, -1 on over-subscription (build_precode_table
 * marker_inflate.rs:4515-4540 rejects code > 2^bits). Incomplete codes are
 * accepted; unused codewords fail when decoded. */
static int huff_build(huff_t *h, const uint8_t *len, int n) {
    uint16_t offs[16];
    memset(h->count, 0, sizeof(h->count));
    for (int s = 0; s < n; s++) h->count[len[s]]++;
    h->count[0] = 0;
    int left = 1;
    for (int l = 1; l < 16; l++) {
        left <<= 1;
        left -= h->count[l];
        if (left < 0) return -1;
    }
    offs[1] = 0;
    for (int l = 1; l < 15; l++) offs[l + 1] = offs[l] + h->count[l];
    for (int s = 0; s < n; s++)
        if (len[s]) h->sorted[offs[len[s]]++] = (uint16_t)s;
    memset(h->lut, 0, sizeof(h->lut));
    /* canonical codes; cache those of length <= HROOT bit-reversed */
    uint32_t code = 0;
    int idx = 0;
    for (int l = 1; l < 16; l++) {
        for (int k = 0; k < h->count[l]; k++, idx++, code++) {
            if (l > HROOT) continue;
            uint32_t rev = 0;
            for (int b = 0; b < l; b++) rev |= ((code >> b) & 1u) << (l - 1 - b);
            for (uint32_t f = rev; f < (1u << HROOT); f += 1u << l)
                h->lut[f] = (uint16_t)(h->sorted[idx] << 4 | l);
        }
        code <<= 1;
    }
    return 0;
}

/* Decode one symbol; returns symbol or -1.  Peeks 15 bits (zero past end). */
static inline int huff_decode(const huff_t *h, bitrd *b) {
    uint32_t bits = (uint32_t)br_peek(b, 15);
    uint16_t e = h->lut[bits & ((1u << HROOT) - 1)];
    if (e & 15) {
        b->pos += e & 15;
        return e >> 4;
    }
    /* canonical walk, one bit at a time (MSB of the code first) */
    int code = 0, first = 0, index = 0;
    for (int l = 1; l < 16; l++) {
        code |= (bits >> (l - 1)) & 1;
        int cnt = h->count[l];
        if (code - cnt < first) {
            b->pos += l;
            return h->sorted[index + (code - first)];
        }
        index += cnt;
        first += cnt;
        first <<= 1;
        code <<= 1;
    }
    return -1;
}

/* Length / distance tables, RFC 1951 3.2.5 (marker_inflate.rs:3818-3835). */
static const uint16_t LBASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                   31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
static const uint8_t LEXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                   2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
static const uint16_t DBASE[30] = {1,    2,    3,    4,    5,    7,     9,     13,    17,  25,
                                   33,   49,   65,   97,   129,  193,   257,   385,   513, 769,
                                   1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
static const uint8_t DEXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2,  3,  3,  4,  4,  5,  5,  6,
                                   6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
/* precode order (RFC 1951 3.2.7; PRECODE_ALPHABET) */
static const uint8_t PRE_ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

/* read_dynamic_huffman_coding (marker_inflate.rs:1206-1245) and
 * read_literal_and_distance_code_lengths (:4447-4510). */
static int read_dynamic_header(bitrd *b, huff_t *lit, huff_t *dist) {
    uint8_t pre[19] = {0}, lens[320];
    int hlit = (int)br_read(b, 5) + 257;
    int hdist = (int)br_read(b, 5) + 1;
    int hclen = (int)br_read(b, 4) + 4;
    if (hlit > 286 || hdist > 30) return -1;
    for (int i = 0; i < hclen; i++) pre[PRE_ORDER[i]] = (uint8_t)br_read(b, 3);
    huff_t ph;
    if (huff_build(&ph, pre, 19)) return -1;
    int total = hlit + hdist;
    for (int i = 0; i < total;) {
        int sym = huff_decode(&ph, b);
        if (sym < 0) return -1;
        if (sym < 16) {
            lens[i++] = (uint8_t)sym;
        } else {
            int rep, val = 0;
            if (sym == 16) {
                if (i == 0) return -1;
                val = lens[i - 1];
                rep = 3 + (int)br_read(b, 2);
            } else if (sym == 17) {
                rep = 3 + (int)br_read(b, 3);
            } else {
                rep = 11 + (int)br_read(b, 7);
            }
            if (i + rep > total) return -1;
            while (rep--) lens[i++] = (uint8_t)val;
        }
    }
    if (lens[256] == 0) return -1;
    if (huff_build(lit, lens, hlit)) return -1;
    if (huff_build(dist, lens + hlit, hdist)) return -1;
    return 0;
}

static void fixed_tables(huff_t *lit, huff_t *dist) {
    uint8_t l[288];
    int i = 0;
    for (; i < 144; i++) l[i] = 8;
    for (; i < 256; i++) l[i] = 9;
    for (; i < 280; i++) l[i] = 7;
    for (; i < 288; i++) l[i] = 8;
    huff_build(lit, l, 288);
    for (i = 0; i < 30; i++) l[i] = 5;
    huff_build(dist, l, 30);
}

/* ===================================================================== */
/* Inflate loop.  One body, two element types (u8 clean / u16 marker).    */
/* Decode semantics: marker_inflate.rs Block::read (:1251) and            */
/* decode_clean_into_contig (:2871); marker values per init_marker_zone   */
/* (:255-268); stop rule per chunk_decode.rs:2029-2040.                   */
/* ===================================================================== */

#define INFLATE_BODY(T, OUT, WINVAL)                                                             \
    do {                                                                                         \
        T *out = (OUT);                                                                          \
        for (;;) {                                                                               \
            uint64_t hbit = b.pos;                                                               \
            if (hbit + 3 > total_bits) goto fail;                                                \
            uint32_t hdr = br_read(&b, 3);                                                       \
            uint32_t bfinal = hdr & 1, btype = hdr >> 1;                                         \
            if (btype == 3) goto fail;                                                           \
            if (hbit >= stop_bit && !bfinal) {                                                   \
                res->end_bit = hbit;                                                             \
                goto done;                                                                       \
            }                                                                                    \
            if (nb < blocks_cap) {                                                               \
                blocks[nb].bit = hbit;                                                           \
                blocks[nb].out_off = op;                                                         \
                blocks[nb].btype = btype;                                                        \
                blocks[nb].bfinal = bfinal;                                                      \
            }                                                                                    \
            nb++;                                                                                \
            if (btype == 0) {                                                                    \
                b.pos = (b.pos + 7) & ~7ull;                                                     \
                if (b.pos + 32 > total_bits) goto fail;                                          \
                uint32_t len = br_read(&b, 16), nlen = br_read(&b, 16);                          \
                if ((len ^ 0xFFFF) != nlen) goto fail;                                           \
                if (b.pos + 8ull * len > total_bits) goto fail;                                  \
                if (op + len > out_cap) goto full;                                               \
                const uint8_t *src = in + (b.pos >> 3);                                          \
                for (uint32_t i = 0; i < len; i++) out[op + i] = (T)src[i];                      \
                op += len;                                                                       \
                b.pos += 8ull * len;                                                             \
            } else {                                                                             \
                const huff_t *lh = &fl, *dh = &fd;                                               \
                if (btype == 2) {                                                                \
                    if (read_dynamic_header(&b, &dl, &dd)) goto fail;                            \
                    lh = &dl;                                                                    \
                    dh = &dd;                                                                    \
                }                                                                                \
                for (;;) {                                                                       \
                    int sym = huff_decode(lh, &b);                                               \
                    if (sym < 0) goto fail;                                                      \
                    if (sym < 256) {                                                             \
                        if (op >= out_cap) goto full;                                            \
                        out[op++] = (T)sym;                                                      \
                    } else if (sym == 256) {                                                     \
                        break;                                                                   \
                    } else {                                                                     \
                        sym -= 257;                                                              \
                        if (sym >= 29) goto fail;                                                \
                        uint32_t len = LBASE[sym] + br_read(&b, LEXTRA[sym]);                    \
                        int ds = huff_decode(dh, &b);                                            \
                        if (ds < 0 || ds >= 30) goto fail;                                       \
                        uint64_t dist = DBASE[ds] + br_read(&b, DEXTRA[ds]);                     \
                        if (dist > op + 32768) goto fail;                                        \
                        if (op + len > out_cap) goto full;                                       \
                        for (uint32_t i = 0; i < len; i++, op++) {                               \
                            if (dist > op) { /* reaches into the window */                       \
                                uint64_t wi = 32768 + op - dist;                                 \
                                out[op] = (T)(WINVAL);                                           \
                            } else {                                                             \
                                out[op] = out[op - dist];                                        \
                            }                                                                    \
                        }                                                                        \
                    }                                                                            \
                    if (b.pos > total_bits) goto fail;                                           \
                }                                                                                \
            }                                                                                    \
            if (b.pos > total_bits) goto fail;                                                   \
            if (bfinal) {                                                                        \
                res->bfinal_hit = 1;                                                             \
                res->end_bit = b.pos;                                                            \
                goto done;                                                                       \
            }                                                                                    \
        }                                                                                        \
    } while (0)

int gzo_inflate(const uint8_t *in, size_t in_len, uint64_t start_bit, uint64_t stop_bit,
                int marker_mode, const uint8_t *window, uint8_t *out8, uint16_t *out16,
                size_t out_cap, gzo_block *blocks, size_t blocks_cap, gzo_result *res) {
    huff_t fl, fd, dl, dd;
    fixed_tables(&fl, &fd);
    bitrd b = {in, in_len, start_bit};
    uint64_t total_bits = (uint64_t)in_len * 8;
    uint64_t op = 0, nb = 0;
    memset(res, 0, sizeof(*res));
    if (marker_mode) {
        INFLATE_BODY(uint16_t, out16, 32768 + wi);
    } else {
        INFLATE_BODY(uint8_t, out8, window ? window[wi] : 0);
    }
done:
    res->out_len = op;
    res->n_blocks = nb;
    res->status = GZO_OK;
    return GZO_OK;
fail:
    res->end_bit = b.pos;
    res->out_len = op;
    res->n_blocks = nb;
    res->status = GZO_DECODE_FAILED;
    return GZO_DECODE_FAILED;
full:
    res->end_bit = b.pos;
    res->out_len = op;
    res->n_blocks = nb;
    res->status = GZO_OUTPUT_FULL;
    return GZO_OUTPUT_FULL;
}

int gzo_resolve_markers(const uint16_t *in, size_t n, const uint8_t *window, uint8_t *out) {
    for (size_t i = 0; i < n; i++) {
        uint16_t v = in[i];
        if (v < 256)
            out[i] = (uint8_t)v;
        else if (v >= 32768)
            out[i] = window[v - 32768];
        else
            return GZO_DECODE_FAILED; /* 256 <= v < 32768 is invalid */
    }
    return GZO_OK;
}

/* ===================================================================== */
/* Block finder (blockfinder_validation.rs)                               */
/* ===================================================================== */

/* validate_precode (:190-229) + parse_precode/is_valid_huffman_lengths
 * (:772-834), restated per definition: precode lengths are 3-bit values; the
 * Kraft leaf count at depth 7 must be 128, or 64 with no length > 1. */
static int precode_ok(const uint8_t *lens19) {
    uint32_t leaves = 0, gt1 = 0;
    for (int i = 0; i < 19; i++) {
        if (lens19[i]) {
            leaves += 1u << (7 - lens19[i]);
            if (lens19[i] > 1) gt1++;
        }
    }
    if (leaves == 64) return gt1 == 0;
    return leaves == 128;
}

/* is_valid_huffman_lengths (:807-834): Kraft equality at 15 bits, or the
 * single-symbol tree (sum 2^14 with no length > 1). */
static int lengths_ok(const uint8_t *l, int n) {
    uint32_t sum = 0, gt1 = 0;
    for (int i = 0; i < n; i++) {
        if (!l[i]) continue;
        sum += 1u << (15 - l[i]);
        if (l[i] > 1) gt1++;
    }
    if (sum == (1u << 14)) return gt1 == 0;
    return sum == (1u << 15);
}

/* Full dynamic-header predicate at one bit offset: is_deflate_candidate_n on
 * 13 bits (:68-98), the precode checks, then validate_huffman_codes
 * (:836-944).  Bits past the end read as zero; `reader.is_eof()` (:131)
 * fails a code-length read that starts at or past the end; the scalar scan
 * breaks when bit + 17 reaches the end (:693-696). */
int gzo_is_dynamic_candidate(const uint8_t *in, size_t in_len, uint64_t bit) {
    uint64_t total = (uint64_t)in_len * 8;
    if (bit + 17 >= total) return 0;
    bitrd b = {in, in_len, bit};
    uint32_t h = br_read(&b, 17);
    if ((h & 1) != 0) return 0;               /* BFINAL must be 0 */
    if (((h >> 1) & 3) != 2) return 0;         /* BTYPE must be 10 */
    uint32_t hlit = (h >> 3) & 31, hdist = (h >> 8) & 31, hclen = (h >> 13) & 15;
    if (hlit > 29 || hdist > 29) return 0;
    uint8_t pre[19] = {0};
    for (uint32_t i = 0; i < hclen + 4; i++) pre[PRE_ORDER[i]] = (uint8_t)br_read(&b, 3);
    if (!precode_ok(pre)) return 0;
    /* precode table (7-bit) */
    huff_t ph;
    if (huff_build(&ph, pre, 19)) return 0;
    int hlit_count = 257 + (int)hlit;
    int total_codes = hlit_count + 1 + (int)hdist;
    uint8_t lens[320];
    for (int i = 0; i < total_codes;) {
        if (b.pos >= total) return 0; /* is_eof */
        int sym = huff_decode(&ph, &b);
        if (sym < 0) return 0;
        if (sym < 16) {
            lens[i++] = (uint8_t)sym;
        } else if (sym == 16) {
            if (i == 0) return 0;
            int rep = 3 + (int)br_read(&b, 2);
            uint8_t prev = lens[i - 1];
            for (int k = 0; k < rep; k++) {
                if (i >= total_codes) return 0;
                lens[i++] = prev;
            }
        } else {
            int rep = sym == 17 ? 3 + (int)br_read(&b, 3) : 11 + (int)br_read(&b, 7);
            for (int k = 0; k < rep; k++) {
                if (i >= total_codes) return 0;
                lens[i++] = 0;
            }
        }
    }
    if (lens[256] == 0) return 0;
    return lengths_ok(lens, hlit_count) && lengths_ok(lens + hlit_count, total_codes - hlit_count);
}

static int cmp_boundary(const void *a, const void *b) {
    const gzo_boundary *x = a, *y = b;
    if (x->bit_offset != y->bit_offset) return x->bit_offset < y->bit_offset ? -1 : 1;
    if (x->seek_bit != y->seek_bit) return x->seek_bit < y->seek_bit ? -1 : 1;
    return 0;
}

/* find_blocks (:520-527) = find_dynamic_blocks (:536) ∪
 * find_uncompressed_blocks (:370-470), sorted by (bit_offset, seek_bit). */
size_t gzo_find_blocks(const uint8_t *in, size_t n, uint64_t start_bit, uint64_t end_bit,
                       gzo_boundary *out, size_t cap) {
    size_t cnt = 0;
    uint64_t total = (uint64_t)n * 8;
    if (end_bit > total) end_bit = total;
    for (uint64_t bit = start_bit; bit < end_bit; bit++) {
        if (gzo_is_dynamic_candidate(in, n, bit)) {
            if (cnt < cap) {
                bitrd b = {in, n, bit + 3};
                out[cnt].bit_offset = bit;
                out[cnt].seek_bit = bit;
                out[cnt].kind = 2;
                out[cnt].hlit = (uint8_t)br_read(&b, 5);
                out[cnt].hdist = (uint8_t)br_read(&b, 5);
                out[cnt].hclen = (uint8_t)br_read(&b, 4);
                out[cnt].pad = 0;
            }
            cnt++;
        }
    }
    /* stored: LEN == ~NLEN at byte b; 3 zero header bits before b; earliest
     * start = 8b - trailing zero bits (max 10); seek = 8b - 3. */
    if (n >= 4) {
        uint64_t until = (end_bit + 7) / 8;
        if (until > n - 4) until = n - 4;
        uint64_t sb = (start_bit + 3 + 7) / 8;
        if (sb < 1) sb = 1;
        for (uint64_t by = sb; by < until; by++) {
            uint32_t size = (uint32_t)in[by] | (uint32_t)in[by + 1] << 8 |
                            (uint32_t)in[by + 2] << 16 | (uint32_t)in[by + 3] << 24;
            if (((size ^ (size >> 16)) & 0xFFFF) != 0xFFFF) continue;
            uint64_t byte_bit = by * 8;
            if (byte_bit < 10) continue;
            uint64_t first = byte_bit - 10;
            uint32_t raw = (uint32_t)in[first / 8] | (uint32_t)in[first / 8 + 1] << 8;
            uint32_t pre = (raw >> (first % 8)) & 0x3FF;
            if (pre & (7u << 7)) continue;
            unsigned tz = 3;
            for (unsigned j = 4; j <= 10; j++) {
                if (pre & (1u << (10 - j))) break;
                tz = j;
            }
            uint64_t earliest = byte_bit - tz, latest = byte_bit - 3;
            if (!(earliest >= start_bit && latest < end_bit)) continue;
            uint32_t len = size & 0xFFFF;
            if (len == 0 || by + 4 + len > n) continue;
            if (cnt < cap) {
                out[cnt].bit_offset = earliest;
                out[cnt].seek_bit = latest;
                out[cnt].kind = 0;
                out[cnt].hlit = out[cnt].hdist = out[cnt].hclen = out[cnt].pad = 0;
            }
            cnt++;
        }
    }
    if (cnt <= cap) qsort(out, cnt, sizeof(gzo_boundary), cmp_boundary);
    return cnt;
}

/* ===================================================================== */
/* Whole-stream serial decode with member walk and trailer verification   */
/* (sm_driver.rs:84-224; mod.rs:725-741 trailing-garbage clean stop)      */
/* ===================================================================== */

int gzo_gunzip(const uint8_t *in, size_t n, uint8_t *out, size_t out_cap, size_t *out_len,
               size_t *members) {
    size_t off = 0, op = 0, m = 0;
    *out_len = 0;
    if (members) *members = 0;
    if (n < 2 || in[0] != 0x1f || in[1] != 0x8b) return GZO_OK; /* non-gzip: Ok(0) */
    for (;;) {
        size_t hl;
        int st = gzo_parse_header(in + off, n - off, &hl);
        if (st) {
            if (m > 0) break; /* trailing bytes after a valid member: clean stop */
            return st;
        }
        gzo_result r;
        const uint8_t *body = in + off + hl;
        size_t body_len = n - off - hl;
        st = gzo_inflate(body, body_len, 0, UINT64_MAX, 0, NULL, out + op, NULL, out_cap - op,
                         NULL, 0, &r);
        if (st) {
            *out_len = op + r.out_len;
            return st == GZO_OUTPUT_FULL ? GZO_OUTPUT_FULL : GZO_DECODE_FAILED;
        }
        size_t end_byte = (size_t)((r.end_bit + 7) / 8);
        if (end_byte + 8 > body_len) {
            *out_len = op + r.out_len;
            return GZO_DECODE_FAILED;
        }
        const uint8_t *t = body + end_byte;
        uint32_t crc = (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16 | (uint32_t)t[3] << 24;
        uint32_t isz = (uint32_t)t[4] | (uint32_t)t[5] << 8 | (uint32_t)t[6] << 16 | (uint32_t)t[7] << 24;
        if ((uint32_t)r.out_len != isz) {
            *out_len = op + r.out_len;
            return GZO_SIZE_MISMATCH;
        }
        if (gzo_crc32(0, out + op, r.out_len) != crc) {
            *out_len = op + r.out_len;
            return GZO_CRC_MISMATCH;
        }
        op += r.out_len;
        m++;
        off += hl + end_byte + 8;
        if (off + 2 > n || in[off] != 0x1f || in[off + 1] != 0x8b) break;
    }
    *out_len = op;
    if (members) *members = m;
    return GZO_OK;
}
