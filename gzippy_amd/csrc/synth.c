/*
 * synth.c — seeded synthetic content generators for the decompression
 * workloads (plain C, host only).
 *
 *   kind 0  text:   order-1 Markov chain over 96 printable bytes (0x20-0x7E
 *                   and '\n'); each row is Zipf(s=1.2) over a per-row random
 *                   permutation.  At each position, with probability 0.35, a
 *                   run of length U[3,258] is copied instead from a
 *                   back-distance floor(exp(U*ln 32768)) (log-uniform in
 *                   1..32768), clamped to the bytes emitted so far.
 *   kind 1  random: uniform random bytes.
 *   kind 2  runs:   byte runs of geometric length (mean 200), run byte uniform.
 *   kind 3  u32:    little-endian uint32 sequence whose deltas are uniform
 *                   in [-8, 8] (wrapping).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} xo_t;

static uint64_t splitmix(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static void xo_seed(xo_t *r, uint64_t seed) {
    for (int i = 0; i < 4; i++) r->s[i] = splitmix(&seed);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t xo_next(xo_t *r) {
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}
static inline double xo_unit(xo_t *r) { return (double)(xo_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static inline uint32_t xo_below(xo_t *r, uint32_t n) { return (uint32_t)(((xo_next(r) >> 32) * (uint64_t)n) >> 32); }

#define NSYM 96

static void gen_text(uint8_t *out, uint64_t n, xo_t *r) {
    uint8_t alpha[NSYM];
    for (int i = 0; i < 95; i++) alpha[i] = (uint8_t)(0x20 + i);
    alpha[95] = '\n';
    /* Zipf(s=1.2) CDF over ranks 1..96 */
    double cdf[NSYM], acc = 0;
    for (int k = 0; k < NSYM; k++) acc += pow((double)(k + 1), -1.2), cdf[k] = acc;
    for (int k = 0; k < NSYM; k++) cdf[k] /= acc;
    /* per-row permutation: perm[row][rank] = symbol index */
    uint8_t perm[NSYM][NSYM];
    for (int row = 0; row < NSYM; row++) {
        for (int k = 0; k < NSYM; k++) perm[row][k] = (uint8_t)k;
        for (int k = NSYM - 1; k > 0; k--) {
            int j = (int)xo_below(r, (uint32_t)k + 1);
            uint8_t t = perm[row][k];
            perm[row][k] = perm[row][j];
            perm[row][j] = t;
        }
    }
    uint8_t sym_of[256];
    memset(sym_of, 0, sizeof(sym_of));
    for (int i = 0; i < NSYM; i++) sym_of[alpha[i]] = (uint8_t)i;
    const double ln32k = log(32768.0);
    uint64_t p = 0;
    int prev = 0;
    while (p < n) {
        if (p > 0 && xo_unit(r) < 0.35) {
            uint64_t len = 3 + xo_below(r, 256);
            uint64_t dist = (uint64_t)floor(exp(xo_unit(r) * ln32k));
            if (dist < 1) dist = 1;
            if (dist > 32768) dist = 32768;
            if (dist > p) dist = p;
            if (len > n - p) len = n - p;
            for (uint64_t i = 0; i < len; i++, p++) out[p] = out[p - dist];
            prev = sym_of[out[p - 1]];
            continue;
        }
        double u = xo_unit(r);
        int lo = 0, hi = NSYM - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (u < cdf[mid]) hi = mid; else lo = mid + 1;
        }
        int s = perm[prev][lo];
        out[p++] = alpha[s];
        prev = s;
    }
}

static void gen_random(uint8_t *out, uint64_t n, xo_t *r) {
    uint64_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t v = xo_next(r);
        memcpy(out + i, &v, 8);
    }
    if (i < n) {
        uint64_t v = xo_next(r);
        memcpy(out + i, &v, n - i);
    }
}

static void gen_runs(uint8_t *out, uint64_t n, xo_t *r) {
    const double lq = log(1.0 - 1.0 / 200.0);
    uint64_t p = 0;
    while (p < n) {
        double u = xo_unit(r);
        if (u <= 0) u = 1e-300;
        uint64_t len = 1 + (uint64_t)floor(log(u) / lq);
        uint8_t b = (uint8_t)xo_next(r);
        if (len > n - p) len = n - p;
        memset(out + p, b, len);
        p += len;
    }
}

static void gen_u32(uint8_t *out, uint64_t n, xo_t *r) {
    uint32_t x = (uint32_t)xo_next(r);
    uint64_t p = 0;
    while (p < n) {
        x += (uint32_t)((int32_t)xo_below(r, 17) - 8);
        uint8_t b[4] = {(uint8_t)x, (uint8_t)(x >> 8), (uint8_t)(x >> 16), (uint8_t)(x >> 24)};
        uint64_t k = n - p < 4 ? n - p : 4;
        memcpy(out + p, b, k);
        p += k;
    }
}

/* Fill out[0..n) with `kind` content from `seed`. Returns 0, or -1 for a bad kind. */
int gzs_generate(int kind, uint8_t *out, uint64_t n, uint64_t seed) {
    xo_t r;
    xo_seed(&r, seed);
    switch (kind) {
    case 0: gen_text(out, n, &r); return 0;
    case 1: gen_random(out, n, &r); return 0;
    case 2: gen_runs(out, n, &r); return 0;
    case 3: gen_u32(out, n, &r); return 0;
    default: return -1;
    }
}
