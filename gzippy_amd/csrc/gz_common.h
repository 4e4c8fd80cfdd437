// gz_common.h — shared device definitions for the gfx950 DEFLATE kernels.
//
// Layout conventions
//   * The compressed input is one byte buffer in HBM; every bit position is
//     absolute (bit b = bit (b & 7) of byte b >> 3, LSB first, RFC 1951 3.1.1).
//     Reads past the buffer end return zero bits (bit_reader.rs:46-133 reads
//     zero-padded bits the same way).
//   * A workgroup of the inflate kernel is ONE wavefront (64 lanes).  The
//     decoder state (bit buffer, positions, counters) is wave-uniform; LDS
//     lookups use a uniform address (broadcast, no bank conflicts) and are
//     pulled back to SGPRs with readfirstlane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gz {

constexpr int WAVE = 64;
constexpr uint32_t WIN = 32768;            // DEFLATE window (RFC 1951 3.2)
constexpr uint32_t MARKER_BASE = 32768;     // replace_markers.rs:23
constexpr int PRE_ROOT = 7;                 // precode LUT covers every precode length

// Piece status (device → host)
enum : uint32_t {
    PS_NONE = 0,
    PS_STOPPED = 1,   // reached a non-final block header at or after stop_bit
    PS_FINAL = 2,     // decoded the final (BFINAL) block
    PS_OVERFLOW = 3,  // slab / output capacity exhausted
    PS_ERROR = 4,     // invalid stream from this start
};
// error sub-codes (PieceOut::err)
enum : uint32_t {
    E_NONE = 0, E_BTYPE3, E_TRUNC, E_STORED_LEN, E_HDR_COUNTS, E_PRECODE, E_CODELEN, E_EOB_ZERO,
    E_LL_TABLE, E_D_TABLE, E_BAD_SYM, E_BAD_DIST_SYM, E_DIST_FAR, E_SYNC,
};

// LUT entry (u32): [0,4) code length (0 = resolve on the slow path),
// [4,8) extra bits, [8,10) kind, [16,32) value.
enum : uint32_t { K_LIT = 0, K_LEN = 1, K_EOB = 2, K_BAD = 3 };

__constant__ uint16_t c_lbase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                     31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
__constant__ uint8_t c_lextra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                     2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
__constant__ uint16_t c_dbase[30] = {1,    2,    3,    4,    5,    7,     9,     13,    17,  25,
                                     33,   49,   65,   97,   129,  193,   257,   385,   513, 769,
                                     1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
__constant__ uint8_t c_dextra[30] = {0, 0, 0, 0, 1, 1, 2, 2,  3,  3,  4,  4,  5,  5,  6,
                                     6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
// precode symbol order (RFC 1951 3.2.7)
__constant__ uint8_t c_pre_order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (uint64_t)hi << 32 | lo;
}
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
// The i1 form: the wave mask comes straight from the compare that computes
// p (HIP's int-typed __ballot widens p to a VGPR and compares it again).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
// This lane's bit of the wave-uniform mask m (the inverse of ballot): used
// as a select condition it is the SGPR mask itself, with no per-lane shift.
__device__ __forceinline__ bool lane_in(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
// d < 0 as a signed 64-bit value, tested on the high word (an SALU compare;
// LLVM lowers a 64-bit signed compare to the VALU).
__device__ __forceinline__ bool neg64(uint64_t d) { return (int32_t)(uint32_t)(d >> 32) < 0; }
// number of active lanes below this one in mask
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ void compiler_fence() { __asm__ __volatile__("" ::: "memory"); }

// Inclusive prefix sum over the 64 lanes: Hillis-Steele inside each row of
// 16 lanes (DPP row_shr 1, 2, 4, 8; lanes shifted in from outside the row
// read 0), then the row totals are carried forward with the GFX9 row
// broadcasts (row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2
// and 3).
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// Pointers that reach a kernel through a struct are generic; accesses through
// them compile to FLAT instructions, which count against lgkmcnt as well as
// vmcnt, so every later LDS wait would also wait for outstanding HBM stores.
// Casting to the global address space keeps them GLOBAL (vmcnt only).
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T *p) {
    return (gptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<const T> to_global(const T *p) {
    return (gptr<const T>)p;
}

// 32-bit little-endian word `wi` of the input; bytes past n read as zero.
__device__ __forceinline__ uint32_t load_word(const uint8_t *__restrict__ in_, uint64_t n, uint64_t wi) {
    gptr<const uint8_t> in = to_global(in_);
    uint64_t b = wi * 4;
    if (b + 4 <= n) return ((gptr<const uint32_t>)in)[wi];
    uint32_t v = 0;
    for (int i = 0; i < 4; i++)
        if (b + i < n) v |= (uint32_t)in[b + i] << (8 * i);
    return v;
}

// 64 bits starting at absolute bit `bit` (zero past the end).
__device__ __forceinline__ uint64_t peek64(const uint8_t *__restrict__ in, uint64_t n, uint64_t bit) {
    uint64_t wi = bit >> 5;
    uint32_t sh = (uint32_t)(bit & 31);
    uint64_t lo = (uint64_t)load_word(in, n, wi) | (uint64_t)load_word(in, n, wi + 1) << 32;
    uint64_t hi = load_word(in, n, wi + 2);
    uint64_t v = lo >> sh;
    if (sh) v |= hi << (64 - sh);
    return v;
}

// ---- wave-uniform bit reader -------------------------------------------
// Holds 33..64 valid bits after refill(); position = next_word*32 - cnt.
// The input is staged through the wave's registers: lane j holds input words
// wbase + 4j .. wbase + 4j + 3 (1 KiB per wave, one coalesced 16-byte load
// per lane).  A refill is a readlane; the window is reloaded synchronously
// once per 1 KiB.  No load is left in flight across decode iterations: on
// this ISA vmcnt also counts stores, so a wait for a prefetch would also wait
// for every output store still on its way to HBM.
struct BitReader {
    const uint8_t *in;
    uint64_t n;
    uint64_t buf;
    uint64_t next_word;
    uint64_t wbase;  // multiple of 4
    uint32_t w0, w1, w2, w3;
    uint32_t cnt;

    __device__ __forceinline__ void load_window(uint64_t base) {
        wbase = base;
        const uint64_t wi = base + 4 * (uint64_t)lane_id();
        if ((wi + 4) * 4 <= n && ((reinterpret_cast<uintptr_t>(in) & 15) == 0)) {
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            u32x4 v = ((gptr<const u32x4>)to_global(in))[wi >> 2];
            w0 = v.x;
            w1 = v.y;
            w2 = v.z;
            w3 = v.w;
        } else {
            w0 = load_word(in, n, wi);
            w1 = load_word(in, n, wi + 1);
            w2 = load_word(in, n, wi + 2);
            w3 = load_word(in, n, wi + 3);
        }
        // wait here (vmcnt(0); expcnt and lgkmcnt left at their maxima) so no
        // load is pending where this path merges with the common refill path
        __builtin_amdgcn_s_waitcnt(0x0F70);
    }
    __device__ __forceinline__ uint32_t word_at(uint64_t wi) {
        uint64_t k = wi - wbase;
        if (k >= 256) {
            load_window(wi & ~3ull);
            k = wi - wbase;
        }
        const int l = (int)(k >> 2);
        const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)w0, l);
        const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)w1, l);
        const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)w2, l);
        const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)w3, l);
        const uint32_t r = (uint32_t)k & 3;
        return r == 0 ? a : r == 1 ? b : r == 2 ? c : d;
    }
    __device__ __forceinline__ void init(const uint8_t *in_, uint64_t n_, uint64_t bit) {
        in = in_;
        n = n_;
        uint64_t wi = bit >> 5;
        uint32_t sh = (uint32_t)(bit & 31);
        load_window(wi & ~3ull);
        uint64_t lo = word_at(wi), hi = word_at(wi + 1);
        buf = (lo | hi << 32) >> sh;
        cnt = 64 - sh;
        next_word = wi + 2;
    }
    __device__ __forceinline__ void refill() {
        if (cnt <= 32) {
            buf |= (uint64_t)word_at(next_word) << cnt;
            next_word++;
            cnt += 32;
        }
    }
    __device__ __forceinline__ uint32_t peek(uint32_t k) const { return (uint32_t)(buf & ((1ull << k) - 1)); }
    __device__ __forceinline__ void consume(uint32_t k) {
        buf >>= k;
        cnt -= k;
    }
    __device__ __forceinline__ uint32_t read(uint32_t k) {
        uint32_t v = peek(k);
        consume(k);
        return v;
    }
    __device__ __forceinline__ uint64_t pos() const { return next_word * 32 - cnt; }
    // re-assert that the reader state is wave-uniform (keeps it in SGPRs)
    __device__ __forceinline__ void make_uniform() {
        buf = uni64(buf);
        cnt = uni(cnt);
        next_word = uni64(next_word);
        wbase = uni64(wbase);
    }
};

__device__ __forceinline__ uint32_t make_ll_entry(uint32_t sym, uint32_t len) {
    if (sym < 256) return len | (K_LIT << 8) | (sym << 16);
    if (sym == 256) return len | (K_EOB << 8);
    if (sym < 286) {
        uint32_t i = sym - 257;
        return len | ((uint32_t)c_lextra[i] << 4) | (K_LEN << 8) | ((uint32_t)c_lbase[i] << 16);
    }
    return len | (K_BAD << 8);
}
__device__ __forceinline__ uint32_t make_d_entry(uint32_t sym, uint32_t len) {
    if (sym < 30) return len | ((uint32_t)c_dextra[sym] << 4) | (K_LEN << 8) | ((uint32_t)c_dbase[sym] << 16);
    return len | (K_BAD << 8);
}

// Canonical decode (RFC 1951 3.2.2) of one codeword from `bits`, MSB of the
// code first; `count` = codes per length.  Used for codes longer than the
// LUT root.  Returns the symbol and its length, or -1.
__device__ __forceinline__ int slow_decode(uint32_t bits, const uint16_t *count, const uint16_t *sorted,
                                           uint32_t &len) {
    int code = 0, first = 0, index = 0;
#pragma unroll
    for (int l = 1; l <= 15; l++) {
        code |= (bits >> (l - 1)) & 1;
        int c = count[l];
        if (code - c < first) {
            len = (uint32_t)l;
            return sorted[index + (code - first)];
        }
        index += c;
        first += c;
        first <<= 1;
        code <<= 1;
    }
    return -1;
}

}  // namespace gz
