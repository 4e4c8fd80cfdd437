// k_inflate.h — speculative chunk inflater: one wavefront per piece.
//
// A "piece" is a bit range of one DEFLATE stream that starts at a block
// header (known from the stream start or from the block finder) and ends at
// the first non-final block header at or after its stop bit, or at the end of
// the final block (stop rule of chunk_decode.rs:2029-2040).
//
// The most recent entries of the sliding window live in LDS as a ring (4096
// u16 or 8192 u8 entries); copies read the ring and write both the ring and
// the piece's output.  A copy from farther back (up to the 32768 of the
// window) reads the piece's own output from HBM, or the window before the
// piece.
//
// Window / output modes
//   DIRECT : the window is all zeros (stream start, WindowMap{0: zeros},
//            chunk_fetcher.rs:694-696); bytes go straight to the member's
//            final output.
//   WINDOW : the caller supplies the 32 KiB window; bytes to dst.
//   MARKER : window unknown.  The ring starts with the marker values
//            32768 + index-from-oldest (replace_markers.rs:6-23;
//            init_marker_zone, marker_inflate.rs:255-268), so a copy that
//            reaches before the piece start copies markers for free.  Output
//            is u16 entries (< 256 literal, >= 32768 marker) until the last
//            32768 entries hold no marker (WidthRing::should_flip,
//            width_ring.rs:310); from then on the output is u8.
//
// Two kernels: the u16-ring kernel runs every piece of a call that has
// MARKER pieces (DIRECT and WINDOW pieces too, so that one launch holds the
// whole call), the u8-ring kernel calls without them.  Both use 10 KB of LDS
// per wavefront, 16 wavefronts per CU (four per SIMD).
//
// Slab layout of a MARKER piece (bytes): [u16 entries: 2*n16] (256-aligned)
// [u8 entries: n8].
#pragma once
#include "gz_common.h"

namespace gz {

// MIDBLOCK: the piece starts inside the block whose header is at hdr_bit, at
// the sync point of start_bit (sync_point).  STOPMID: the piece stops at the
// sync point of stop_bit inside the block whose header is at stop_hdr, when
// its decode lands there exactly; else at the first non-final block header
// at or after stop_bit, as without the flag.
enum : uint32_t { PF_DIRECT = 1, PF_MARKER = 2, PF_WINDOW = 4, PF_MIDBLOCK = 8, PF_STOPMID = 16 };

struct PieceDesc {
    uint64_t start_bit;      // first block header to decode
    uint64_t stop_bit;       // stop before a non-final block header at >= stop_bit
    uint64_t limit_bit;      // end of the member's readable deflate bits
    uint8_t *dst;            // DIRECT/WINDOW: output bytes; MARKER: slab base
    uint64_t cap;            // DIRECT/WINDOW: output capacity (bytes); MARKER: slab bytes
    const uint8_t *window;   // WINDOW: 32768 bytes, oldest first
    uint32_t flags;
    uint32_t member;
    uint64_t hdr_bit;        // MIDBLOCK: header of the block holding start_bit
    uint64_t stop_hdr;       // STOPMID: header of the block holding stop_bit
    float weight;            // issue priority in a launch with more pieces than wavefronts: the
                             // piece's planned cost over the launch's largest (0: none)
    uint32_t pad_w;
};

struct PieceOut {
    uint64_t start_bit;  // first symbol decoded (MIDBLOCK: the sync point), else start_bit
    uint64_t end_bit;
    uint64_t n16;       // u16 marker-mode entries
    uint64_t n8;        // u8 entries (DIRECT/WINDOW, or after the flip)
    uint64_t u8_off;    // slab byte offset of the u8 entries
    uint64_t bits_done; // bits consumed when the piece stopped (sizing on overflow)
    uint64_t t_begin, t_end;  // wall clock (100 MHz) around the piece
    uint32_t status;
    uint32_t err;
    uint32_t nblocks;
    uint32_t mode;
    // the first block header the piece decodes at or after its start (NO_CP:
    // none), and the piece's entries before it: a point where a decode can
    // resume with the 32 KiB of output before it as the window
    uint64_t cp_bit;
    uint64_t cp_out;
};
constexpr uint64_t NO_CP = ~0ull;

// Ring of the most recent entries in LDS: 2048 u16 or 4096 u8 entries, 4 KiB
// either way, so that with the LUTs a wavefront needs 10 KB and 16 share a
// CU, four per SIMD (the kernel stays at 128 VGPRs: the dynamic-header
// reader, whose table builds are the register peak, is not inlined), which
// hides the decode loop's LDS, scalar-load and readlane latencies.
// A copy from farther back (up to the 32768 of the window) reads the piece's
// own output from HBM, or the window before the piece.
template <typename RT>
constexpr uint32_t RING = sizeof(RT) == 2 ? 2048u : 4096u;
template <typename RT>
constexpr uint32_t RMASK = RING<RT> - 1;

// Issue-priority balance state of a wavefront (progress_report), in LDS: it
// is read once every PROG_EVERY rounds and would otherwise hold scalar
// registers through the whole decode loop.
struct ProgState {
    uint64_t *ctr;         // launch-wide progress (null: by remaining work, with `weight`)
    uint64_t start, last;  // piece's first bit, position at the last report
    float inv_total, inv_span, weight;
};
typedef __attribute__((address_space(3))) ProgState lds_prog;

// LW: the litlen LUT root width (9 or 10 bits), a kernel instantiation.
template <typename RT, uint32_t LW = 9>
struct __align__(16) InflateShared {
    // Two-level Huffman LUTs: a root of LLR / DR bits, then subtables for
    // the codes longer than the root (zlib's inflate_table layout), in the
    // same array after the root.  9/8-bit roots keep a wavefront's LDS at
    // 10 KB (16 wavefronts per CU); a 10-bit litlen root takes 1.8 KB more
    // (13 per CU) and sends fewer codes to a subtable.  The subtable reserve
    // covers the worst case of zlib's bound (ENOUGH_LENS: 852 entries for a
    // 9-bit root, 1334 for 10); a subtable that does not fit its reserve
    // leaves its codes to the canonical slow path.
    static constexpr int LLR = (int)LW;
    static constexpr int DR = 8;
    static constexpr int LL_SUB = LW == 9 ? 384 : 320;
    static constexpr int D_SUB = 64;
    RT ring[RING<RT>];
    uint32_t ll[(1 << LLR) + LL_SUB];
    uint32_t dl[(1 << DR) + D_SUB];
    uint16_t ll_sorted[288];
    uint16_t d_sorted[32];
    uint16_t ll_count[16];
    uint16_t d_count[16];
    uint16_t offs[16];
    uint16_t firstc[16];
    uint8_t lens[320];
    uint8_t pre_lens[32];
    uint32_t qi;
    uint32_t sub_next;
    // far-copy sources: piece positions [0, far_n16) are u16 entries at
    // far_out16, from far_n16 on u8 entries at far_out8 (far_n16 moves from
    // "all" to the flip position when a marker piece flips); before position
    // 0 the window (far_win_kind 0: none, 1: u8 buffer, 3: markers)
    uint64_t far_out16, far_out8, far_n16, far_win;
    uint32_t far_win_kind;
    ProgState pg;
};

// ---- canonical table construction (whole wave) ----------------------------
// MODE 0: litlen, 1: distance, 2: precode (entry = len | sym << 16).
// Returns 0, or -1 when the lengths over-subscribe the code space
// (build_precode_table, marker_inflate.rs:4515-4540).  Incomplete codes are
// accepted; their unused codewords fail when decoded.
template <int MODE, typename SH>
__device__ int build_table(SH &sh, const uint8_t *lens, int n, uint32_t *lut, int R, int sub_cap, uint16_t *count,
                           uint16_t *sorted) {
    const uint32_t lane = lane_id();
    uint32_t l[5], rank[5];
#pragma unroll
    for (int c = 0; c < 5; c++) {
        int s = c * WAVE + (int)lane;
        l[c] = s < n ? lens[s] : 0u;
        rank[c] = 0;
    }
    uint32_t run[16];
#pragma unroll
    for (int L = 0; L < 16; L++) run[L] = 0;
#pragma unroll
    for (int c = 0; c < 5; c++) {
        if (c * WAVE < n) {
#pragma unroll
            for (int L = 1; L <= 15; L++) {
                uint64_t m = ballot(l[c] == (uint32_t)L);
                if (l[c] == (uint32_t)L) rank[c] = run[L] + mbcnt(m);
                run[L] += (uint32_t)__popcll(m);
            }
        }
    }
    int left = 1;
    bool over = false;
#pragma unroll
    for (int L = 1; L <= 15; L++) {
        left = 2 * left - (int)run[L];
        over |= left < 0;
    }
    if (over) return -1;
    const bool complete = left == 0;
    uint32_t offs = 0, code = 0;
#pragma unroll
    for (int L = 1; L <= 15; L++) {
        code = (code + run[L - 1]) << 1;
        if (lane == (uint32_t)L) {
            sh.offs[L] = (uint16_t)offs;
            sh.firstc[L] = (uint16_t)code;
            count[L] = (uint16_t)run[L];
        }
        offs += run[L];
    }
    if (lane == 0) count[0] = 0;
    if (!complete)
        for (uint32_t i = lane; i < (1u << R); i += WAVE) lut[i] = 0;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 5; c++) {
        uint32_t s = c * WAVE + lane;
        uint32_t L = l[c];
        if ((int)s < n && L) {
            uint32_t cd = (uint32_t)sh.firstc[L] + rank[c];
            sorted[sh.offs[L] + rank[c]] = (uint16_t)s;
            uint32_t rev = __brev(cd) >> (32 - L);
            if ((int)L <= R) {
                uint32_t e = MODE == 0 ? make_ll_entry(s, L) : MODE == 1 ? make_d_entry(s, L) : (L | (s << 16));
                for (uint32_t f = rev; f < (1u << R); f += 1u << L) lut[f] = e;
            } else {
                lut[rev & ((1u << R) - 1)] = 0;
            }
        }
    }
    __syncthreads();
    // ---- second level for codes longer than the root.  Canonical codes
    // sharing their first R bits are consecutive in canonical order; each
    // such group gets a subtable of 2^(longest - R) entries, indexed by the
    // next bits of the stream.  Root entry of a group: length 0, extra field
    // = subtable bits, value = subtable index.
    uint32_t nlong = 0;
#pragma unroll
    for (int L = 1; L <= 15; L++) nlong += L > R ? run[L] : 0u;
    if (sub_cap > 0 && nlong > 0) {
        if (lane == 0) {
            uint32_t next = 1u << R, gprefix = ~0u, glen = 0;
            auto close_group = [&]() {
                const uint32_t bits = glen - (uint32_t)R;
                if (next + (1u << bits) <= (1u << R) + (uint32_t)sub_cap) {
                    lut[__brev(gprefix) >> (32 - R)] = (next << 16) | (bits << 4);
                    next += 1u << bits;
                }
            };
            for (uint32_t L = (uint32_t)R + 1; L <= 15; L++) {
                const uint32_t cntL = count[L], fc = sh.firstc[L];
                for (uint32_t i = 0; i < cntL; i++) {
                    const uint32_t prefix = (fc + i) >> (L - (uint32_t)R);
                    if (prefix != gprefix) {
                        if (gprefix != ~0u) close_group();
                        gprefix = prefix;
                    }
                    glen = L;
                }
            }
            close_group();
            sh.sub_next = next;
        }
        __syncthreads();
        if (!complete)  // unused codewords of an incomplete code must read as invalid
            for (uint32_t i = (1u << R) + lane; i < sh.sub_next; i += WAVE) lut[i] = 0;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 5; c++) {
            uint32_t s = c * WAVE + lane;
            uint32_t L = l[c];
            if ((int)s < n && (int)L > R) {
                const uint32_t code = (uint32_t)sh.firstc[L] + rank[c];
                const uint32_t k = L - (uint32_t)R;
                const uint32_t pe = lut[__brev(code >> k) >> (32 - R)];
                if ((pe & 15) == 0 && (pe >> 16) != 0) {
                    const uint32_t sb = (pe >> 4) & 15, off = pe >> 16;
                    const uint32_t sidx = __brev(code & ((1u << k) - 1)) >> (32 - k);
                    const uint32_t e = MODE == 0 ? make_ll_entry(s, L) : make_d_entry(s, L);
                    for (uint32_t f = sidx; f < (1u << sb); f += 1u << k) lut[off + f] = e;
                }
            }
        }
        __syncthreads();
    }
    return 0;
}

// Two-level lookup: the root entry, or the subtable entry it points to.
__device__ __forceinline__ uint32_t sub_lookup(const uint32_t *lut, uint32_t e, uint64_t bits_after_root) {
    if ((e & 15) == 0 && (e >> 16) != 0)
        e = lut[(e >> 16) + ((uint32_t)bits_after_root & ((1u << ((e >> 4) & 15)) - 1))];
    return e;
}

template <typename SH>
__device__ void build_fixed(SH &sh) {
    const uint32_t lane = lane_id();
    for (uint32_t s = lane; s < 288; s += WAVE) sh.lens[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8;
    __syncthreads();
    build_table<0>(sh, sh.lens, 288, sh.ll, SH::LLR, SH::LL_SUB, sh.ll_count, sh.ll_sorted);
    for (uint32_t s = lane; s < 32; s += WAVE) sh.lens[s] = 5;
    __syncthreads();
    build_table<1>(sh, sh.lens, 32, sh.dl, SH::DR, SH::D_SUB, sh.d_count, sh.d_sorted);
}

// read_dynamic_huffman_coding (marker_inflate.rs:1206-1245) and
// read_literal_and_distance_code_lengths (:4447-4510).  Returns E_NONE or an
// error code.
template <typename SH>
__device__ __attribute__((noinline)) uint32_t read_dynamic(SH &sh, BitReader &br) {
    const uint32_t lane = lane_id();
    br.refill();
    uint32_t hlit = br.read(5) + 257, hdist = br.read(5) + 1, hclen = br.read(4) + 4;
    if (hlit > 286 || hdist > 30) return E_HDR_COUNTS;
    // up to 19 x 3 = 57 precode bits, read as <= 30 + <= 27 (a refill
    // guarantees 33 bits)
    const uint32_t n_lo = hclen < 10 ? hclen : 10, n_hi = hclen - n_lo;
    br.refill();
    uint64_t p = br.read(3 * n_lo);
    br.refill();
    p |= (uint64_t)br.read(3 * n_hi) << (3 * n_lo);
    // lane i holds the length of the i-th transmitted precode
    uint32_t pl = lane < hclen ? (uint32_t)(p >> (3 * lane)) & 7 : 0;
    for (uint32_t s = lane; s < 32; s += WAVE) sh.pre_lens[s] = 0;
    __syncthreads();
    if (lane < hclen) sh.pre_lens[c_pre_order[lane]] = (uint8_t)pl;
    __syncthreads();
    // the precode LUT lives in the litlen LUT storage until the litlen table is built
    if (build_table<2>(sh, sh.pre_lens, 19, sh.ll, PRE_ROOT, 0, sh.d_count, sh.d_sorted)) return E_PRECODE;
    const uint32_t total = hlit + hdist;
    uint32_t i = 0, prev = 0;
    while (i < total) {
        br.refill();
        uint32_t e = uni(sh.ll[br.peek(PRE_ROOT)]);
        uint32_t L = e & 15;
        if (L == 0) return E_CODELEN;
        uint32_t sym = e >> 16;
        br.consume(L);
        if (sym < 16) {
            if (lane == 0) sh.lens[i] = (uint8_t)sym;
            prev = sym;
            i++;
        } else {
            uint32_t rep, val;
            if (sym == 16) {
                if (i == 0) return E_CODELEN;
                val = prev;
                rep = 3 + br.read(2);
            } else if (sym == 17) {
                val = 0;
                rep = 3 + br.read(3);
            } else {
                val = 0;
                rep = 11 + br.read(7);
            }
            if (i + rep > total) return E_CODELEN;
            for (uint32_t k = lane; k < rep; k += WAVE) sh.lens[i + k] = (uint8_t)val;
            prev = val;
            i += rep;
        }
    }
    __syncthreads();
    // LDS values that steer control flow are read back through readfirstlane
    // so the compiler keeps the branch (and everything after it) uniform
    if (uni(sh.lens[256]) == 0) return E_EOB_ZERO;
    if (build_table<0>(sh, sh.lens, (int)hlit, sh.ll, SH::LLR, SH::LL_SUB, sh.ll_count, sh.ll_sorted)) return E_LL_TABLE;
    if (build_table<1>(sh, sh.lens + hlit, (int)hdist, sh.dl, SH::DR, SH::D_SUB, sh.d_count, sh.d_sorted)) return E_D_TABLE;
    return E_NONE;
}

// ---- output segment ----------------------------------------------------
// Entries go to `p` as u16 (wide: a marker piece before its flip, only in
// the u16-ring kernel) or u8.  `seg0` is the piece position of entry 0 of
// this segment (the ring index of entry i is seg0 + i).
struct OutSeg {
    gptr<uint8_t> p;
    uint64_t pos, cap, seg0;
    uint32_t wide;
};

template <typename RT>
__device__ __forceinline__ bool is_wide(const OutSeg &o) {
    return sizeof(RT) == 2 && o.wide;
}

// Entries are u8 in the u8-ring kernel at compile time; the u16-ring kernel
// writes u16 until its flip.
template <typename RT>
__device__ __forceinline__ void put(const OutSeg &o, uint64_t i, uint32_t v) {
    if (is_wide<RT>(o)) ((gptr<uint16_t>)o.p)[i] = (uint16_t)v;
    else o.p[i] = (uint8_t)v;
}
// Entry soff + voff of the segment (soff wave-uniform, voff per lane) as a
// 32-bit byte offset from the segment base: the store takes the base from
// SGPRs and the offset from one VGPR, with no 64-bit address arithmetic.
// Segment positions stay below 2^30 + a round, so offsets fit 32 bits.
template <typename RT>
__device__ __forceinline__ void put_e(const OutSeg &o, uint32_t soff, uint32_t voff, uint32_t v) {
    if (is_wide<RT>(o)) *(gptr<uint16_t>)(o.p + (uint64_t)((soff + voff) * 2u)) = (uint16_t)v;
    else o.p[(uint64_t)(soff + voff)] = (uint8_t)v;
}
// The ring slot at byte offset b (taken modulo the ring's size in bytes).
// Callers pass a wave-uniform part plus a per-lane part, both in bytes, so
// a slot address is one add and one AND.
template <typename RT>
__device__ __forceinline__ RT &ring_at(RT *ring, uint32_t b) {
    return *reinterpret_cast<RT *>(reinterpret_cast<uint8_t *>(ring) + (b & (RING<RT> * (uint32_t)sizeof(RT) - 1u)));
}
// A use of v on every path (an empty asm statement taking it as an
// operand).  Values loaded from HBM are otherwise used only inside
// lane-masked branches, which the compiler's wait-count analysis must assume
// can be skipped; a load it believes may still be in flight at the top of
// the decode loop costs a full vmcnt wait (on this ISA, for every store
// before it too) each time its register is reused.
__device__ __forceinline__ void consumed(uint32_t v) { __asm__ __volatile__("" ::"v"(v)); }
// Lanes of a chunk inside the copy: the first `rem` lanes (rem >= 1).
__device__ __forceinline__ uint64_t first_lanes(uint32_t rem) { return rem >= WAVE ? ~0ull : (1ull << rem) - 1; }
// Lanes whose entry is a marker (>= 256), as a wave mask straight from the
// compare (no intermediate boolean).
__device__ __forceinline__ uint64_t marker_lanes(uint32_t v) { return __builtin_amdgcn_uicmp(v, 255u, 34 /* unsigned > */); }

// ---- LZ77 copy by the whole wave -----------------------------------------
// Entry pos+i takes the value at pos - dist + (i mod dist): every source lies
// before pos, so no entry of a copy depends on another entry of the same
// copy.  All reads (at most five chunks of 64 for a 258-byte copy) are issued
// before any write, with one wait in between; the only ring slot both read
// and written is dist == ring size for the same entry, read first.  LDS
// operations of one wave execute in order, which orders a copy after the
// symbols before it.  All index math is 32-bit (segment positions stay
// below 2^31).
//
// The output width W (1: u16 entries, 0: u8) is a template parameter of the
// copy, decided once per copy; chunks are written through copy_chunk, every
// chunk but the copy's last without a lane mask.

// Marker bookkeeping of a copy while the output is u16.  last_marker_end is
// a bound: every entry from it to the current position is a literal (the
// flip waits until it is a window behind).  A copy whose source starts at or
// after it copies literals only (an overlapping copy repeats its own source)
// and needs no check; otherwise a marker among the copied entries moves the
// bound to the copy's end (at most 257 entries past the exact last marker,
// which only delays the flip by as much).
struct MarkerTrack {
    bool on;     // the copy may copy markers
    uint64_t m;  // lanes of the chunks so far that copied one
};
template <uint32_t W>
__device__ __forceinline__ MarkerTrack marker_track(const OutSeg &o, uint32_t pos, uint32_t dist,
                                                    uint64_t last_marker_end) {
    return {W != 0 && (int64_t)(o.seg0 + pos) - (int64_t)dist < (int64_t)last_marker_end, 0ull};
}
__device__ __forceinline__ void marker_done(const MarkerTrack &t, const OutSeg &o, uint32_t pos, uint32_t len,
                                            uint64_t &last_marker_end) {
    if (t.m) last_marker_end = o.seg0 + pos + len;
}

// Entry soff + voff of the segment, width W known at compile time (see put_e).
template <uint32_t W>
__device__ __forceinline__ void put_w(const OutSeg &o, uint32_t soff, uint32_t voff, uint32_t v) {
    if (W) *(gptr<uint16_t>)(o.p + (uint64_t)(soff * 2u + voff * 2u)) = (uint16_t)v;
    else o.p[(uint64_t)(soff + voff)] = (uint8_t)v;
}

// Chunk kb (entries kb .. kb + 63 of the copy, lane j holding entry kb + j)
// to the ring and the output.  FULL: the copy covers the whole chunk (no lane
// mask); otherwise lanes past the copy's end write nothing.
template <typename RT, uint32_t W, bool FULL>
__device__ __forceinline__ void copy_chunk(RT *ring, const OutSeg &o, uint32_t gpos, uint32_t pos, uint32_t kb,
                                           uint32_t len, uint32_t v, MarkerTrack &t) {
    const uint32_t lane = lane_id();
    constexpr uint32_t SZ = (uint32_t)sizeof(RT);
    if (FULL || lane + kb < len) {
        ring_at<RT>(ring, (gpos + kb) * SZ + lane * SZ) = (RT)v;
        put_w<W>(o, pos + kb, lane, v);
    }
    if (W && t.on) t.m |= FULL ? marker_lanes(v) : marker_lanes(v) & first_lanes(len - kb);
    consumed(v);
}

// Where entries more than a ring back are read from (see InflateShared).
struct FarSrc {
    uint64_t out16, out8, n16, win;
    uint32_t kind;
};
template <typename SH>
__device__ __forceinline__ FarSrc far_src(SH &sh) {
    return {uni64(sh.far_out16), uni64(sh.far_out8), uni64(sh.far_n16), uni64(sh.far_win), uni(sh.far_win_kind)};
}
// The entry at piece position r (negative: the window before the piece).
// Entries more than 32768 before the piece do not exist (a distance the
// window cannot serve): they read as 0 without a memory access.
template <typename RT>
__device__ __forceinline__ uint32_t far_fetch(const FarSrc &f, int64_t r) {
    if (r >= 0) {
        if (sizeof(RT) == 2 && (uint64_t)r < f.n16) return to_global((const uint16_t *)f.out16)[r];
        return to_global((const uint8_t *)f.out8)[(uint64_t)r - f.n16];
    }
    if (r >= -(int64_t)WIN && f.kind == 1) return to_global((const uint8_t *)f.win)[WIN + r];
    if (r >= -(int64_t)WIN && f.kind == 3) return MARKER_BASE + (uint32_t)(WIN + r);
    return 0u;
}

// A copy whose source is more than a ring back, in NC chunks: the piece's
// output in HBM, written by this wavefront at least a ring earlier, or the
// window before the piece (never overlapping the copy: dist > ring > len).
// A wavefront's vector memory operations reach the CU's L1 in order and the
// loads' wait covers its earlier stores.  Each chunk count has its own
// straight-line body, so every load issued is consumed on the same path:
// no load is left in flight past the copy, where the decode loop would have
// to wait for it (and, on this ISA, for every store before it).
template <typename RT, uint32_t W, uint32_t NC>
__device__ __forceinline__ void far_store(RT *ring, const OutSeg &o, uint32_t pos, uint32_t gpos, uint32_t len,
                                          uint32_t dist, const uint32_t (&v)[NC], uint64_t &last_marker_end) {
    MarkerTrack t = marker_track<W>(o, pos, dist, last_marker_end);
#pragma unroll
    for (uint32_t k = 0; k + 1 < NC; k++) copy_chunk<RT, W, true>(ring, o, gpos, pos, WAVE * k, len, v[k], t);
    copy_chunk<RT, W, false>(ring, o, gpos, pos, WAVE * (NC - 1), len, v[NC - 1], t);
    marker_done(t, o, pos, len, last_marker_end);
    compiler_fence();
}
template <typename RT, uint32_t W, uint32_t NC>
__device__ __forceinline__ void far_copy(RT *ring, const OutSeg &o, const FarSrc &f, uint32_t pos, uint32_t gpos,
                                         uint32_t len, uint32_t dist, uint64_t &last_marker_end) {
    const uint32_t lane = lane_id();
    const int64_t rel = (int64_t)o.seg0 + (int64_t)pos - (int64_t)dist;  // piece position of entry 0
    uint32_t v[NC];
#pragma unroll
    for (uint32_t k = 0; k < NC; k++) {
        const uint32_t i = lane + WAVE * k;
        v[k] = i < len ? far_fetch<RT>(f, rel + (int64_t)i) : 0u;
    }
    compiler_fence();
    far_store<RT, W, NC>(ring, o, pos, gpos, len, dist, v, last_marker_end);
}

template <typename RT, uint32_t W, typename SH>
__device__ __attribute__((always_inline)) void wave_copy_w(SH &sh, const OutSeg &o, uint32_t pos, uint32_t g0,
                                                           uint32_t len, uint32_t dist, uint64_t &last_marker_end) {
    RT *ring = sh.ring;
    const uint32_t lane = lane_id();
    const uint32_t gpos = g0 + pos;
    constexpr uint32_t SZ = (uint32_t)sizeof(RT);
    if (dist > RING<RT>) {
        const FarSrc f = far_src(sh);
        const uint32_t nc = (len + WAVE - 1) / WAVE;
        if (nc <= 1) far_copy<RT, W, 1>(ring, o, f, pos, gpos, len, dist, last_marker_end);
        else if (nc == 2) far_copy<RT, W, 2>(ring, o, f, pos, gpos, len, dist, last_marker_end);
        else if (nc == 3) far_copy<RT, W, 3>(ring, o, f, pos, gpos, len, dist, last_marker_end);
        else if (nc == 4) far_copy<RT, W, 4>(ring, o, f, pos, gpos, len, dist, last_marker_end);
        else far_copy<RT, W, 5>(ring, o, f, pos, gpos, len, dist, last_marker_end);
        return;
    }
    const uint32_t sb = (gpos - dist) * SZ;  // ring byte of the source's first entry (before wrapping)
    // every read before any write (ring offsets wrap, so lanes past the
    // copy's end read harmlessly); chunk k's value is read only when the copy
    // reaches it
    uint32_t v0, v1, v2, v3, v4;
    if (dist < len) {
        // The copy overlaps itself: entry i reads source entry i mod dist.
        // Lane j's remainder comes from the hardware reciprocal (exact except
        // at multiples of dist, where the quotient can come out one short and
        // the remainder equal to dist, folded to 0 by the unsigned min);
        // each next chunk adds 64 mod dist (lane 63's remainder + 1) and
        // folds once, since both terms are below dist.
        const float fd = (float)dist;
        const float q = __builtin_floorf((float)lane * __builtin_amdgcn_rcpf(fd));
        const uint32_t r = (uint32_t)__builtin_fmaf(-q, fd, (float)lane);
        uint32_t s = min(r, r - dist);
        const uint32_t r63 = uni((uint32_t)__builtin_amdgcn_readlane((int)s, 63));
        const uint32_t step = r63 + 1 == dist ? 0u : r63 + 1;
        auto adv = [&](uint32_t x) {
            const uint32_t t = x + step;
            return min(t, t - dist);
        };
        v0 = ring_at<RT>(ring, sb + s * SZ);
        if (len > 64) {
            s = adv(s);
            v1 = ring_at<RT>(ring, sb + s * SZ);
            if (len > 128) {
                s = adv(s);
                v2 = ring_at<RT>(ring, sb + s * SZ);
                if (len > 192) {
                    s = adv(s);
                    v3 = ring_at<RT>(ring, sb + s * SZ);
                    if (len > 256) {
                        s = adv(s);
                        v4 = ring_at<RT>(ring, sb + s * SZ);
                    }
                }
            }
        }
    } else {
        const uint32_t lb = lane * SZ;
        v0 = ring_at<RT>(ring, sb + lb);
        if (len > 64) {
            v1 = ring_at<RT>(ring, sb + 64 * SZ + lb);
            if (len > 128) {
                v2 = ring_at<RT>(ring, sb + 128 * SZ + lb);
                if (len > 192) {
                    v3 = ring_at<RT>(ring, sb + 192 * SZ + lb);
                    if (len > 256) v4 = ring_at<RT>(ring, sb + 256 * SZ + lb);
                }
            }
        }
    }
    compiler_fence();
    // chunk k is full when the copy reaches past it (len > 64 (k + 1))
    MarkerTrack t = marker_track<W>(o, pos, dist, last_marker_end);
    if (len <= 64) {
        copy_chunk<RT, W, false>(ring, o, gpos, pos, 0, len, v0, t);
    } else {
        copy_chunk<RT, W, true>(ring, o, gpos, pos, 0, len, v0, t);
        if (len <= 128) {
            copy_chunk<RT, W, false>(ring, o, gpos, pos, 64, len, v1, t);
        } else {
            copy_chunk<RT, W, true>(ring, o, gpos, pos, 64, len, v1, t);
            if (len <= 192) {
                copy_chunk<RT, W, false>(ring, o, gpos, pos, 128, len, v2, t);
            } else {
                copy_chunk<RT, W, true>(ring, o, gpos, pos, 128, len, v2, t);
                if (len <= 256) {
                    copy_chunk<RT, W, false>(ring, o, gpos, pos, 192, len, v3, t);
                } else {
                    copy_chunk<RT, W, true>(ring, o, gpos, pos, 192, len, v3, t);
                    copy_chunk<RT, W, false>(ring, o, gpos, pos, 256, len, v4, t);
                }
            }
        }
    }
    marker_done(t, o, pos, len, last_marker_end);
    compiler_fence();
}

template <typename RT, typename SH>
__device__ __attribute__((always_inline)) void wave_copy(SH &sh, const OutSeg &o, uint32_t pos, uint32_t g0,
                                                         uint32_t len, uint32_t dist, uint64_t &last_marker_end) {
    if (is_wide<RT>(o)) wave_copy_w<RT, 1>(sh, o, pos, g0, len, dist, last_marker_end);
    else wave_copy_w<RT, 0>(sh, o, pos, g0, len, dist, last_marker_end);
}

// Packed symbol of a lookahead lane: a literal is SYM_LIT | its value, a
// copy is distance | length << 16 (distance <= 32768, length <= 258), end-of-
// block is SYM_EOB and an invalid or incomplete code 0.  A complete literal
// or copy is thus any value above SYM_EOB.
constexpr uint32_t SYM_LIT = 0x80000000u, SYM_EOB = 1u;
__device__ __forceinline__ bool sym_is_lit(uint32_t w) { return (int32_t)w < 0; }
// m |= 1 << b for a wave-uniform b, as one scalar instruction
__device__ __forceinline__ void set_bit64(uint64_t &m, uint32_t b) { __asm__("s_bitset1_b64 %0, %1" : "+s"(m) : "s"(b)); }
__device__ __forceinline__ uint32_t sym_len(uint32_t w) { return sym_is_lit(w) ? 1u : w >> 16; }

// A round whose output fits one wavefront (run <= 64 entries) and holds
// copies, in one pass: lane j produces entry j.  Each taken symbol (lanes ta
// of the round's first half, tb of its second, packed as wa / wb, output
// offsets oa / ob) leaves a descriptor (literal value, or copy distance) at
// its first output slot in LDS scratch; a ballot of the slots gives every
// lane its symbol.  A copy
// entry's source j - dist inside the round is chased by pointer jumping
// over ds_bpermute (each step doubles the hop, at most six); a source before
// the round is read from the ring or, more than a ring back, from HBM.  All
// reads precede the round's one ring write and one store, so no cut rule
// applies (sources inside the round never reach the ring or HBM).  Scratch:
// the code-length array, which only the header reader uses.
template <typename RT, typename SH>
__device__ __attribute__((always_inline)) void fused_round(SH &sh, const OutSeg &o, uint32_t pos, uint32_t g0,
                                                           uint32_t run, bool ta, uint32_t wa, uint32_t oa, bool tb,
                                                           uint32_t wb, uint32_t ob, uint64_t &last_marker_end) {
    static_assert(sizeof(sh.lens) >= 4 * WAVE, "scratch");
    const uint32_t lane = lane_id();
    uint32_t *scr = reinterpret_cast<uint32_t *>(sh.lens);
    compiler_fence();
    constexpr uint32_t VALID = 0x80000000u, LIT = 0x40000000u;
    // lanes exchange values through scr: the fences keep the compiler from
    // treating each lane's store and load as its own (forwarding the 0)
    scr[lane] = 0u;
    compiler_fence();
    if (ta) scr[oa] = VALID | (sym_is_lit(wa) ? LIT | (wa & 0xFFu) : (wa & 0xFFFFu));
    if (tb) scr[ob] = VALID | (sym_is_lit(wb) ? LIT | (wb & 0xFFu) : (wb & 0xFFFFu));
    compiler_fence();
    const uint32_t d0 = scr[lane];
    const uint64_t S = ballot(d0 != 0u);  // first output slots of the round's symbols
    const uint64_t below = S & (lane == WAVE - 1 ? ~0ull : (2ull << lane) - 1);
    const uint32_t st = below ? 63u - (uint32_t)__builtin_clzll(below) : 0u;
    const uint32_t d = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(st << 2), (int)d0);
    const bool live = lane < run;
    uint32_t val = 0, ptr = 0;
    bool res = true;
    bool far = false;
    int32_t src = 0;
    if (!(d & LIT)) {
        src = (int32_t)lane - (int32_t)(d & 0xFFFFu);
        if (src >= 0) {
            res = false;
            ptr = (uint32_t)src;
        } else if ((uint32_t)(-src) <= RING<RT>) {
            val = ring_at<RT>(sh.ring, (g0 + pos) * (uint32_t)sizeof(RT) + (uint32_t)src * (uint32_t)sizeof(RT));
        } else {
            far = live;
        }
    } else {
        val = d & 0xFFu;
    }
    if (ballot(far)) {
        const FarSrc f = far_src(sh);
        if (far) val = far_fetch<RT>(f, (int64_t)o.seg0 + (int64_t)pos + (int64_t)src);
    }
    for (uint64_t pend = ballot(!res && live); pend; pend = ballot(!res && live)) {
        const uint32_t w = res ? (VALID | val) : ptr;
        const uint32_t x = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(ptr << 2), (int)w);
        if (!res) {
            if (x & VALID) {
                val = x & 0xFFFFu;
                res = true;
            } else {
                ptr = x;
            }
        }
    }
    compiler_fence();
    if (live) {
        ring_at<RT>(sh.ring, (g0 + pos) * (uint32_t)sizeof(RT) + lane * (uint32_t)sizeof(RT)) = (RT)val;
        put_e<RT>(o, pos, lane, val);
    }
    if (is_wide<RT>(o) && (marker_lanes(val) & first_lanes(run))) last_marker_end = o.seg0 + pos + run;
    consumed(val);
    compiler_fence();
}

enum : uint32_t { R_EOB = 0, R_FLIP = 1, R_OVERFLOW = 2, R_ERROR = 3, R_CONT = 4, R_STOPMID = 5 };

// 128 bits of input starting at absolute bit P, as two wave-uniform words.
// Reads six 32-bit words through the scalar data cache (SMEM loads, counted
// by lgkmcnt — never ordered behind the wave's pending HBM stores); near the
// end of the buffer it falls back to bounds-checked word loads.
__device__ __attribute__((always_inline)) void window128(const uint8_t *__restrict__ in, uint64_t n, uint64_t P,
                                                         uint64_t &W0, uint64_t &W1) {
    const uint64_t wP = P >> 5;
    const uint32_t sh = (uint32_t)(P & 31);
    uint32_t a0, a1, a2, a3, a4, a5;
    if ((wP + 6) * 4 <= n) {
        // the input is never written while the kernel runs: read it through
        // the constant address space so a wave-uniform load becomes SMEM
        typedef __attribute__((address_space(4))) const uint32_t cword;
        cword *w = (cword *)reinterpret_cast<const uint32_t *>(in) + wP;
        a0 = w[0];
        a1 = w[1];
        a2 = w[2];
        a3 = w[3];
        a4 = w[4];
        a5 = w[5];
    } else {
        // uniform too, so both paths merge in scalar registers
        a0 = uni(load_word(in, n, wP));
        a1 = uni(load_word(in, n, wP + 1));
        a2 = uni(load_word(in, n, wP + 2));
        a3 = uni(load_word(in, n, wP + 3));
        a4 = uni(load_word(in, n, wP + 4));
        a5 = uni(load_word(in, n, wP + 5));
    }
    const uint64_t A = uni64((uint64_t)a0 | (uint64_t)a1 << 32), B = uni64((uint64_t)a2 | (uint64_t)a3 << 32),
                   C = uni64((uint64_t)a4 | (uint64_t)a5 << 32);
    W0 = sh ? (A >> sh) | (B << (64 - sh)) : A;
    W1 = sh ? (B >> sh) | (C << (64 - sh)) : B;
}

// 192 bits of input starting at absolute bit P (W0: bits P..P+63, W1 the
// next 64, W2 the 64 after), from seven words read as in window128.
__device__ __attribute__((always_inline)) void window192(const uint8_t *__restrict__ in, uint64_t n, uint64_t P,
                                                         uint64_t &W0, uint64_t &W1, uint64_t &W2) {
    const uint64_t wP = P >> 5;
    const uint32_t sh = (uint32_t)(P & 31);
    uint32_t a[7];
    if ((wP + 7) * 4 <= n) {
        typedef __attribute__((address_space(4))) const uint32_t cword;
        cword *w = (cword *)reinterpret_cast<const uint32_t *>(in) + wP;
#pragma unroll
        for (int k = 0; k < 7; k++) a[k] = w[k];
    } else {
#pragma unroll
        for (int k = 0; k < 7; k++) a[k] = uni(load_word(in, n, wP + (uint64_t)k));
    }
    const uint64_t A = uni64((uint64_t)a[0] | (uint64_t)a[1] << 32), B = uni64((uint64_t)a[2] | (uint64_t)a[3] << 32),
                   C = uni64((uint64_t)a[4] | (uint64_t)a[5] << 32), D = uni64((uint64_t)a[6]);
    W0 = sh ? (A >> sh) | (B << (64 - sh)) : A;
    W1 = sh ? (B >> sh) | (C << (64 - sh)) : B;
    W2 = sh ? (C >> sh) | (D << (64 - sh)) : C;
}

// One symbol at P, the slow way (codes longer than the LUT roots, and any
// condition the multi-symbol round leaves to a single step).
template <typename RT, typename SH>
__device__ __attribute__((always_inline)) uint32_t scalar_step(SH &sh, const uint8_t *__restrict__ in, uint64_t n,
                                                               uint64_t &P, OutSeg &o, uint32_t &pos, uint32_t cap,
                                                               uint32_t g0, uint32_t hist0, uint64_t &last_marker_end,
                                                               uint32_t &err) {
    const uint32_t lane = lane_id();
    uint64_t W0, W1;
    window128(in, n, P, W0, W1);
    uint32_t e = uni(sh.ll[(uint32_t)W0 & ((1u << SH::LLR) - 1)]);
    e = uni(sub_lookup(sh.ll, e, W0 >> SH::LLR));
    if ((e & 15) == 0) {
        uint32_t sl;
        int sy = (int)uni((uint32_t)slow_decode((uint32_t)W0 & 0x7FFF, sh.ll_count, sh.ll_sorted, sl));
        if (sy < 0) {
            err = E_BAD_SYM;
            return R_ERROR;
        }
        e = make_ll_entry((uint32_t)sy, uni(sl));
    }
    e = uni(e);
    const uint32_t L = e & 15, kind = (e >> 8) & 3;
    if (kind == K_BAD) {
        err = E_BAD_SYM;
        return R_ERROR;
    }
    if (kind == K_EOB) {
        P += L;
        return R_EOB;
    }
    if (kind == K_LIT) {
        if (pos >= cap) return R_OVERFLOW;
        if (lane == 0) {
            ring_at<RT>(sh.ring, (g0 + pos) * (uint32_t)sizeof(RT)) = (RT)(e >> 16);
            put_e<RT>(o, pos, 0, e >> 16);
        }
        compiler_fence();
        pos++;
        P += L;
        return R_CONT;
    }
    const uint32_t E = (e >> 4) & 15;
    const uint32_t len = (e >> 16) + ((uint32_t)(W0 >> L) & ((1u << E) - 1));
    if (pos + len > cap) return R_OVERFLOW;
    const uint32_t off = L + E;
    const uint64_t wd = (W0 >> off) | (off ? W1 << (64 - off) : 0);
    uint32_t d = uni(sh.dl[(uint32_t)wd & ((1u << SH::DR) - 1)]);
    d = uni(sub_lookup(sh.dl, d, wd >> SH::DR));
    if ((d & 15) == 0) {
        uint32_t sl;
        int sy = (int)uni((uint32_t)slow_decode((uint32_t)wd & 0x7FFF, sh.d_count, sh.d_sorted, sl));
        if (sy < 0) {
            err = E_BAD_DIST_SYM;
            return R_ERROR;
        }
        d = make_d_entry((uint32_t)sy, uni(sl));
    }
    d = uni(d);
    if (((d >> 8) & 3) != K_LEN) {
        err = E_BAD_DIST_SYM;
        return R_ERROR;
    }
    const uint32_t DL = d & 15, DE = (d >> 4) & 15;
    const uint32_t dist = (d >> 16) + ((uint32_t)(wd >> DL) & ((1u << DE) - 1));
    if (dist > hist0 + pos + WIN) {
        err = E_DIST_FAR;
        return R_ERROR;
    }
    wave_copy<RT>(sh, o, pos, g0, len, dist, last_marker_end);
    pos += len;
    P += off + DL + DE;
    return R_CONT;
}

// The symbols that would start at the first bit of ch[h], h < NH (lane j:
// the input from bit Q_h + j on): the packed form, the lane of the following
// symbol (j + the symbol's bits, at most j + 48), and the wave masks of the
// lanes holding a complete literal or length+distance code (CONT), end-of-
// block (EOBM) and a literal (LITM).  Straight-line per lane: the distance
// code is looked up on every lane (its result kept only where the code is a
// length), and the lane classes are wave masks from the compares themselves.
// The NH decodes run in lockstep, so each LUT level is one LDS round trip
// for all of them.
template <int NH, typename SH>
__device__ __forceinline__ void lane_decode(SH &sh, const uint64_t (&ch)[NH], uint32_t lane, uint32_t (&w)[NH],
                                            uint32_t (&nxt)[NH], uint64_t (&CONT)[NH], uint64_t (&EOBM)[NH],
                                            uint64_t (&LITM)[NH]) {
    constexpr int EQ = 32, NE = 33, UGT = 34;
    uint32_t e[NH], d[NH], off[NH];
    uint64_t SUB = 0, DSUB = 0, SUBh[NH], DSUBh[NH], LENK[NH];
#pragma unroll
    for (int h = 0; h < NH; h++) e[h] = sh.ll[(uint32_t)ch[h] & ((1u << SH::LLR) - 1)];
#pragma unroll
    for (int h = 0; h < NH; h++) {
        SUBh[h] = __builtin_amdgcn_uicmp(e[h] & 15, 0u, EQ) & __builtin_amdgcn_uicmp(e[h], 0xFFFFu, UGT);
        SUB |= SUBh[h];
    }
    // codes longer than the root: one more LDS round trip, only when a lane needs it
    if (SUB) {
#pragma unroll
        for (int h = 0; h < NH; h++)
            if (lane_in(SUBh[h]))
                e[h] = sh.ll[(e[h] >> 16) + ((uint32_t)(ch[h] >> SH::LLR) & ((1u << ((e[h] >> 4) & 15)) - 1))];
    }
#pragma unroll
    for (int h = 0; h < NH; h++) {
        off[h] = (e[h] & 15) + ((e[h] >> 4) & 15);
        d[h] = sh.dl[(uint32_t)(ch[h] >> off[h]) & ((1u << SH::DR) - 1)];
    }
#pragma unroll
    for (int h = 0; h < NH; h++) {
        LENK[h] = __builtin_amdgcn_uicmp((e[h] >> 8) & 3, K_LEN, EQ);
        DSUBh[h] = LENK[h] & __builtin_amdgcn_uicmp(d[h] & 15, 0u, EQ) & __builtin_amdgcn_uicmp(d[h], 0xFFFFu, UGT);
        DSUB |= DSUBh[h];
    }
    if (DSUB) {
#pragma unroll
        for (int h = 0; h < NH; h++)
            if (lane_in(DSUBh[h]))
                d[h] = sh.dl[(d[h] >> 16) +
                             ((uint32_t)(ch[h] >> (off[h] + SH::DR)) & ((1u << ((d[h] >> 4) & 15)) - 1))];
    }
#pragma unroll
    for (int h = 0; h < NH; h++) {
        const uint32_t L = e[h] & 15, kind = (e[h] >> 8) & 3, E = (e[h] >> 4) & 15;
        const uint32_t len = (e[h] >> 16) + ((uint32_t)(ch[h] >> L) & ((1u << E) - 1));
        const uint32_t DL = d[h] & 15, DE = (d[h] >> 4) & 15;
        const uint32_t dist = (d[h] >> 16) + ((uint32_t)(ch[h] >> (off[h] + DL)) & ((1u << DE) - 1));
        const uint64_t NZ = __builtin_amdgcn_uicmp(L, 0u, NE);
        const uint64_t LITK = __builtin_amdgcn_uicmp(kind, K_LIT, EQ), EOBK = __builtin_amdgcn_uicmp(kind, K_EOB, EQ);
        const uint64_t DOK = __builtin_amdgcn_uicmp(DL, 0u, NE) & __builtin_amdgcn_uicmp((d[h] >> 8) & 3, K_LEN, EQ);
        const bool is_len = lane_in(LENK[h]);
        CONT[h] = NZ & (LITK | (LENK[h] & DOK));
        // selects, not branches: every lane computes every form
        const uint32_t sym = is_len ? dist | len << 16 : SYM_LIT | (e[h] >> 16);
        w[h] = lane_in(CONT[h]) ? sym : lane_in(NZ & EOBK) ? SYM_EOB : 0u;
        nxt[h] = lane + (is_len ? off[h] + DL + DE : L);
        EOBM[h] = NZ & EOBK;
        LITM[h] = NZ & LITK;
    }
}

// A round's lookahead: lane j decodes the symbols that would start at bits
// P + j (half A) and P + 64 + j (half B).  A step is the lane of the
// following symbol within the half (up to 111: past the half from 64 on),
// plus STEP_STOP when that lane is inside the half and holds no complete
// literal or copy (a walk stops there).  Lane classes travel in the packed
// symbols and steps rather than in wave masks, which would hold twelve
// scalar registers through the decode loop.
constexpr uint32_t STEP_STOP = 128, STEP_LANE = 127;
struct Look2 {
    uint32_t wa, wb;  // packed symbols
    uint32_t sa, sb;  // steps
};

template <typename SH>
__device__ __attribute__((always_inline)) void lookahead2(SH &sh, const uint8_t *__restrict__ in, uint64_t n,
                                                          uint64_t P, Look2 &la) {
    const uint32_t lane = lane_id();
    uint64_t W0, W1, W2;
    window192(in, n, P, W0, W1, W2);
    const uint64_t cha = lane ? (W0 >> lane) | (W1 << (64 - lane)) : W0;
    const uint64_t chb = lane ? (W1 >> lane) | (W2 << (64 - lane)) : W1;
    const uint64_t ch[2] = {cha, chb};
    uint32_t w[2], x[2];
    uint64_t C[2], E[2], L[2];
    lane_decode<2>(sh, ch, lane, w, x, C, E, L);
    la.wa = w[0];
    la.wb = w[1];
    const bool stop_a = x[0] < 64 && !((C[0] >> (x[0] & 63)) & 1);
    const bool stop_b = x[1] < 64 && !((C[1] >> (x[1] & 63)) & 1);
    la.sa = x[0] + (stop_a ? STEP_STOP : 0u);
    la.sb = x[1] + (stop_b ? STEP_STOP : 0u);
}

// ---- a symbol boundary inside a block, from any bit ---------------------
constexpr uint64_t NO_SYNC = ~0ull;
constexpr uint32_t SYNC_WINDOWS = 64;  // give up after 4096 bits

// The first bit S >= X at which every decode of the current block's codes
// started at X, X+1, ..., X+47 has a symbol boundary, or NO_SYNC.  A symbol
// (code plus extra bits, a length with its distance) spans at most 48 bits,
// so the true symbol sequence has a boundary in [X, X+48) and passes
// through S too.  Paths are followed as a set of pending positions over
// 64-bit lookahead windows; a path at an invalid code or at end-of-block is
// dropped (a start is planned well before the block's end).  The result
// depends only on the block's tables and
// the input, so a piece stopping at X and the piece starting at X find the
// same S; the stopping piece only stops there if its own decode lands on S,
// which is what makes a start at S exact.  Huffman codes resynchronise
// within a few symbols, so S is usually within 100 bits of X.  X below the
// block's first symbol bit `first` is refused.
template <typename SH>
__device__ __attribute__((noinline)) uint64_t sync_point(SH &sh, const uint8_t *__restrict__ in, uint64_t n,
                                                         uint64_t X, uint64_t first, uint64_t limit_bit) {
    X = uni64(X);
    if (X < uni64(first)) return NO_SYNC;
    uint64_t P = X;
    uint64_t A = (1ull << 48) - 1;  // pending path positions, relative to P
    for (uint32_t w = 0; w < SYNC_WINDOWS; w++) {
        if (P + 128 > uni64(limit_bit)) return NO_SYNC;
        uint64_t W0, W1;
        window128(in, n, P, W0, W1);
        const uint32_t lane = lane_id();
        const uint64_t ch[1] = {lane ? (W0 >> lane) | (W1 << (64 - lane)) : W0};
        uint32_t sw[1], nx[1];
        uint64_t CM[1], EM[1], LM[1];
        lane_decode<1>(sh, ch, lane, sw, nx, CM, EM, LM);
        const uint32_t snx = nx[0];
        const uint64_t CONT = uni64(CM[0]);
        uint64_t C = 0;  // pending positions in the next window
        while (A) {
            A = uni64(A);
            C = uni64(C);
            const uint32_t i = uni((uint32_t)__builtin_ctzll(A));
            if (C == 0 && (A & (A - 1)) == 0) return P + i;  // every path is at i
            A &= A - 1;
            // end-of-block or an invalid code: the path stops here.  A
            // misaligned path meets the end-of-block code now and then;
            // the true one only where the block ends, and then the
            // stopping piece does not land on S
            if (!((CONT >> i) & 1)) continue;
            const uint32_t t = uni((uint32_t)__builtin_amdgcn_readlane((int)snx, (int)i));
            if (t < 64) A |= 1ull << t;
            else C |= 1ull << (t - 64);
        }
        A = uni64(C);
        if (A == 0) return NO_SYNC;
        P += 64;
    }
    return NO_SYNC;
}

// Decode the Huffman symbols of the current block until end-of-block, from
// bit P.  Multi-symbol rounds over 128 input bits: lane j looks up the litlen
// (and, for a length code, the distance) LUT entries of the symbols starting
// at bits P + j and P + 64 + j, both halves in one LDS round trip per LUT
// level, and knows where the next symbol would start (j + code bits).  A
// scalar walk from lane 0 of the first half follows those steps through the
// lanes that hold complete codes and on into the second half (about twelve
// symbols per round on text: a round is bound by its input bits), a wave
// prefix sum per half gives every taken symbol its output offset, the round's
// literals go out in one LDS write and one store per half, and its copies run
// in order.  The next round's input window and LUT lookups are issued before
// this round's stores and copies, so their latency overlaps the copies.  A
// round that can take nothing (a code longer than the LUT root, an invalid
// code, a full output) is left to scalar_step.  Against 64-bit rounds the
// walk, the prefix sums and the round's bookkeeping run half as often: on the
// 1 GiB text member 2.52M rounds instead of 4.66M, kernel 3.36 -> 3.23 ms;
// on u32 deltas (bit-bound) 5.87 -> 4.61 ms per 256 MiB.
// ---- balance across the launch: issue priority from progress -------------
// All pieces of a launch are planned to take equal time and, when they fit
// the resident wavefronts, start together,
// but the SIMD's issue arbiter favours its oldest wavefront, so pieces finish
// spread out and the launch waits for the last.  Every PROG_EVERY rounds a
// wavefront adds its piece's progress since the last report (a fraction of
// its piece, in units of 1/PROG_UNIT) to a launch-wide counter (one atomic,
// whose return over the piece count is the launch's mean progress) and sets
// its priority by how far its own fraction is behind that mean: the lagging
// wavefronts issue first.  A scheduling hint only; the output does not
// depend on it.  (Counting consumed bits instead, the launch's progress was
// the share of all bits consumed, which stored pieces, many bits each and
// fast ones, dominate in mixed content: every Huffman piece then looked
// equally behind.)
struct Prog {
    uint32_t on;  // balance on (the state is in ProgState)
    uint32_t k;   // rounds
};
// Measured on a 1 GiB text member (3969 pieces): a report every 64 rounds
// and priority bands 2% of the piece wide took the kernel 3.76 -> 3.61 ms;
// every 16 or 32 rounds the atomics cost more than the balance gains, every
// 128 or 256 balances too late; bands of 0.5% or 5% were slower.
constexpr uint32_t PROG_EVERY = 64;
constexpr float PROG_BAND = 0.02f;
constexpr float PROG_UNIT = 65536.0f;  // counter units per piece

// Where pieces outnumber the resident wavefronts (a batch), they are pulled
// over the launch and fractions are not comparable; every wavefront starts
// at once, so the arbiter's oldest-first order says nothing about which
// piece has the most work left either.  There each piece's priority follows
// its remaining planned work, (1 - its fraction done) x its weight (its cost
// over the launch's largest), in bands: the big pieces, which end the launch
// when they run slow, issue first.
__device__ __forceinline__ void prio_by_remaining(float rem) {
    if (rem > 0.5f) __builtin_amdgcn_s_setprio(3);
    else if (rem > 0.25f) __builtin_amdgcn_s_setprio(2);
    else if (rem > 0.1f) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

__device__ __attribute__((noinline)) void progress_report(lds_prog *ps, uint64_t P) {
    const float inv_span = ps->inv_span;
    if (!ps->ctr) {
        prio_by_remaining((1.0f - (float)(P - uni64(ps->start)) * inv_span) * ps->weight);
        return;
    }
    const uint64_t d = (uint64_t)((float)(P - uni64(ps->last)) * inv_span * PROG_UNIT + 0.5f);
    uint64_t G = 0;
    if (lane_id() == 0) G = atomicAdd((unsigned long long *)ps->ctr, (unsigned long long)d) + d;
    G = uni64(G);
    const float behind = (float)G * ps->inv_total - (float)(P - uni64(ps->start)) * inv_span;
    ps->last = P;
    if (behind > PROG_BAND) __builtin_amdgcn_s_setprio(3);
    else if (behind > 0.0f) __builtin_amdgcn_s_setprio(2);
    else if (behind > -PROG_BAND) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

template <typename RT, typename SH>
__device__ __attribute__((always_inline)) uint32_t decode_symbols(SH &sh, const uint8_t *__restrict__ in, uint64_t n,
                                                                  uint64_t &P, OutSeg &o, uint64_t limit_bit,
                                                                  uint64_t &last_marker_end, uint32_t &err,
                                                                  uint64_t &stop_at, Prog &pg) {
    const uint32_t lane = lane_id();
    // every control value is re-asserted wave-uniform (readfirstlane): the
    // compiler's divergence analysis cannot prove it for values that came
    // through memory, and a "divergent" loop control turns the whole decode
    // loop into exec-masked code with VGPR copies of scalar state
    o.p = (gptr<uint8_t>)uni64((uint64_t)o.p);
        o.pos = uni64(o.pos);
    o.cap = uni64(o.cap);
    o.seg0 = uni64(o.seg0);
    o.wide = uni(o.wide);
    limit_bit = uni64(limit_bit);
    // 32-bit segment positions: the segment is rebased before 2^30 entries
    uint32_t pos = (uint32_t)o.pos;
    uint32_t cap = uni((uint32_t)(o.cap < 0x7FFFFFFFull ? o.cap : 0x7FFFFFFFull));
    uint32_t g0 = uni((uint32_t)o.seg0);
    const uint32_t hist0 = uni((uint32_t)(o.seg0 < WIN ? o.seg0 : WIN));  // history before the segment, capped
    uint32_t r = R_EOB;
    P = uni64(P);
    // nothing loaded before (stored blocks, the header reader) is still in
    // flight when the loop starts; see consumed()
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt unchanged
    Look2 la;
    lookahead2(sh, in, n, P, la);
    for (;;) {
        P = uni64(P);
        pos = uni(pos);
        o.p = (gptr<uint8_t>)uni64((uint64_t)o.p);  // stores then take the base from SGPRs
        if (pg.on && (++pg.k & (PROG_EVERY - 1)) == 0) progress_report((lds_prog *)&sh.pg, P);
        last_marker_end = uni64(last_marker_end);
        stop_at = uni64(stop_at);
        // 64-bit orderings as the sign of a difference (bit positions are
        // below 2^63): SALU subtractions instead of VALU 64-bit compares
        if (stop_at != NO_SYNC && !neg64(P - stop_at)) {
            if (P == stop_at) {
                r = R_STOPMID;
                break;
            }
            stop_at = NO_SYNC;  // the decode went past it (single steps)
        }
        if (neg64(limit_bit - P)) {
            err = E_TRUNC;
            r = R_ERROR;
            break;
        }
        // ---- walk the chain of symbol starts from bit P through half A, and
        // on into half B where it leaves A: each lane's step names the next
        // start.  s ends as the bit (relative to P) after the round's last
        // symbol; an end-of-block symbol ends the walk and the block.
        uint64_t MA = 0, MB = 0, XA = 0, XB = 0;  // symbol starts taken; X*: the end-of-block one
        uint32_t s = 0, eob = 0;
        if (uni((uint32_t)__builtin_amdgcn_readlane((int)la.wa, 0)) > SYM_EOB) {
            do {
                set_bit64(MA, s);
                s = uni((uint32_t)__builtin_amdgcn_readlane((int)la.sa, (int)s));
            } while (s < 64);
        } else {
            s = STEP_STOP;
        }
        if (s >= STEP_STOP) {
            // stopped at lane q of half A: end-of-block, or a code the round
            // cannot take
            const uint32_t q = s - STEP_STOP;
            s = q;
            if (uni((uint32_t)__builtin_amdgcn_readlane((int)la.wa, (int)q)) == SYM_EOB) {
                XA = 1ull << q;
                s = uni((uint32_t)__builtin_amdgcn_readlane((int)la.sa, (int)q)) & STEP_LANE;
                eob = 1;
            }
        } else if (s < 128) {
            uint32_t t = s - 64;
            if (uni((uint32_t)__builtin_amdgcn_readlane((int)la.wb, (int)t)) > SYM_EOB) {
                do {
                    set_bit64(MB, t);
                    t = uni((uint32_t)__builtin_amdgcn_readlane((int)la.sb, (int)t));
                } while (t < 64);
            } else {
                t += STEP_STOP;
            }
            if (t >= STEP_STOP) {
                const uint32_t q = t - STEP_STOP;
                t = q;
                if (uni((uint32_t)__builtin_amdgcn_readlane((int)la.wb, (int)q)) == SYM_EOB) {
                    XB = 1ull << q;
                    t = uni((uint32_t)__builtin_amdgcn_readlane((int)la.sb, (int)q)) & STEP_LANE;
                    eob = 1;
                }
            }
            s = 64 + t;
        }
        MA = uni64(MA | XA);
        MB = uni64(MB | XB);
        // ---- a mid-block stop inside this round: keep the symbols before it
        // when one of them ends exactly there, else the stream does not pass
        // through it and it is dropped
        if (stop_at != NO_SYNC && stop_at - P < (uint64_t)s) {
            const uint32_t rel = uni((uint32_t)(stop_at - P));
            const bool at = rel < 64 ? ((MA >> rel) & 1) != 0 : rel < 128 && ((MB >> (rel - 64)) & 1) != 0;
            if (at) {
                if (rel < 64) {
                    MA &= (1ull << rel) - 1;
                    MB = 0;
                } else {
                    MB &= (1ull << (rel - 64)) - 1;
                }
                s = rel;
                eob = 0;
            } else {
                stop_at = NO_SYNC;
            }
        }
        // ---- output offsets of the taken symbols, half A then half B
        const uint32_t wa = la.wa, wb = la.wb;
        uint64_t TA = uni64(MA & ~XA), TB = uni64(MB & ~XB);  // literals and copies
        const uint32_t va = lane_in(TA) ? sym_len(wa) : 0u;
        const uint32_t vb = lane_in(TB) ? sym_len(wb) : 0u;
        const uint32_t ia = wave_scan_incl(va);
        const uint32_t ib0 = wave_scan_incl(vb);
        const uint32_t runa = uni((uint32_t)__builtin_amdgcn_readlane((int)ia, 63));
        const uint32_t ib = ib0 + runa;
        uint32_t run = runa + uni((uint32_t)__builtin_amdgcn_readlane((int)ib0, 63));
        if ((MA | MB) == 0 || run > cap - pos) {
            // nothing the round can take: one symbol the slow way, which
            // either advances P or ends the block with a status
            const uint32_t sr = scalar_step<RT>(sh, in, n, P, o, pos, cap, g0, hist0, last_marker_end, err);
            if (sr != R_CONT) {
                r = sr;
                break;
            }
            P = uni64(P);
            lookahead2(sh, in, n, P, la);
        } else {
            // The round writes its literals to the ring before its copies, so
            // it is cut short where that order would matter:
            //  - a ring copy whose source slot a later entry of the round
            //    would overwrite first (distance + rest of the round > ring):
            //    the round ends after it;
            //  - a symbol whose output would end more than a ring past the
            //    round's start: a copy's destination slot would alias a
            //    later literal's, and the copy, written second, would win.
            //    The round ends before it;
            //  - a copy from beyond the ring whose source reaches into this
            //    round's own output, not yet in HBM: the round ends before
            //    it.
            // None of these can cut before the round's first symbol (lengths
            // are at most 258, the ring holds 2048 or more entries).
            // A round no longer than the ring can only meet the first: the
            // second needs a symbol ending past the ring, the third a far
            // copy (distance > ring) whose source lies inside the round.
            const uint64_t LITA = ballot(sym_is_lit(wa)), LITB = ballot(sym_is_lit(wb));
            const uint64_t CPA = TA & ~LITA, CPB = TB & ~LITB;
            const uint32_t oa = ia - va, ob = ib - vb;
            const uint32_t da = wa & 0xFFFFu, db = wb & 0xFFFFu;
            uint64_t BADA, BADB, BEFA = 0, BEFB = 0;  // BEF*: the round ends before the symbol
            if (run <= RING<RT>) {
                BADA = CPA & __builtin_amdgcn_uicmp(da, RING<RT>, 37 /* unsigned <= */) &
                       __builtin_amdgcn_uicmp(da + run - oa, RING<RT>, 34 /* unsigned > */);
                BADB = CPB & __builtin_amdgcn_uicmp(db, RING<RT>, 37) & __builtin_amdgcn_uicmp(db + run - ob, RING<RT>, 34);
            } else {
                const bool ca = lane_in(CPA), cb = lane_in(CPB);
                const bool fa = da > RING<RT>, fb = db > RING<RT>;
                BADA = ballot(ca && !fa && da + run - oa > RING<RT>);
                BADB = ballot(cb && !fb && db + run - ob > RING<RT>);
                // the round's first symbol is lane 0 of half A
                BEFA = ballot(lane != 0 && ((ca && fa && da < ia) || (lane_in(TA) && ia > RING<RT>)));
                BEFB = ballot((cb && fb && db < ib) || (lane_in(TB) && ib > RING<RT>));
            }
            if (BADA | BADB | BEFA | BEFB) {  // rare
                // keep the symbols that start before position kp (0..127)
                uint32_t kp = 128;
                if (BADA) kp = uni((uint32_t)__builtin_ctzll(BADA)) + 1;
                else if (BADB) kp = 64 + uni((uint32_t)__builtin_ctzll(BADB)) + 1;
                if (BEFA) {
                    const uint32_t q = uni((uint32_t)__builtin_ctzll(BEFA));
                    kp = q < kp ? q : kp;
                } else if (BEFB) {
                    const uint32_t q = 64 + uni((uint32_t)__builtin_ctzll(BEFB));
                    kp = q < kp ? q : kp;
                }
                const uint64_t nA = kp >= 64 ? MA : MA & ((1ull << kp) - 1);
                const uint64_t nB = kp <= 64 ? 0ull : kp >= 128 ? MB : MB & ((1ull << (kp - 64)) - 1);
                if (nA != MA || nB != MB) {
                    MA = uni64(nA);
                    MB = uni64(nB);
                    if (MB) {
                        const uint32_t last = uni((uint32_t)(63 - __builtin_clzll(MB)));
                        run = uni((uint32_t)__builtin_amdgcn_readlane((int)ib, (int)last));
                        s = 64 + (uni((uint32_t)__builtin_amdgcn_readlane((int)la.sb, (int)last)) & STEP_LANE);
                        TB = uni64(MB);
                    } else {
                        const uint32_t last = uni((uint32_t)(63 - __builtin_clzll(MA)));
                        run = uni((uint32_t)__builtin_amdgcn_readlane((int)ia, (int)last));
                        s = uni((uint32_t)__builtin_amdgcn_readlane((int)la.sa, (int)last)) & STEP_LANE;
                        TB = 0;
                    }
                    TA = uni64(MA);
                    eob = 0;  // the cut is before the end-of-block symbol
                }
            }
            eob = uni(eob);
            const uint64_t lita = uni64(TA & LITA), litb = uni64(TB & LITB);
            uint64_t cpa = uni64(TA & ~LITA), cpb = uni64(TB & ~LITB);
            const bool fused = run <= WAVE && (cpa | cpb);
            P += s;
            if (!eob) lookahead2(sh, in, n, P, la);
            if (fused) {
                fused_round<RT>(sh, o, pos, g0, run, lane_in(TA), wa, oa, lane_in(TB), wb, ob, last_marker_end);
            } else {
                // ---- literals of the round: one LDS write and one HBM store per half
                if (lane_in(lita)) {
                    ring_at<RT>(sh.ring, (g0 + pos) * (uint32_t)sizeof(RT) + oa * (uint32_t)sizeof(RT)) =
                        (RT)(wa & 0xFFu);
                    put_e<RT>(o, pos, oa, wa & 0xFFu);
                }
                if (lane_in(litb)) {
                    ring_at<RT>(sh.ring, (g0 + pos) * (uint32_t)sizeof(RT) + ob * (uint32_t)sizeof(RT)) =
                        (RT)(wb & 0xFFu);
                    put_e<RT>(o, pos, ob, wb & 0xFFu);
                }
                compiler_fence();
                // ---- copies, in stream order
                while (cpa | cpb) {
                    uint32_t m, w;
                    if (cpa) {
                        const uint32_t mi = uni((uint32_t)__builtin_ctzll(cpa));
                        cpa = uni64(cpa & (cpa - 1));
                        m = (uint32_t)__builtin_amdgcn_readlane((int)oa, (int)mi);
                        w = (uint32_t)__builtin_amdgcn_readlane((int)wa, (int)mi);
                    } else {
                        const uint32_t mi = uni((uint32_t)__builtin_ctzll(cpb));
                        cpb = uni64(cpb & (cpb - 1));
                        m = (uint32_t)__builtin_amdgcn_readlane((int)ob, (int)mi);
                        w = (uint32_t)__builtin_amdgcn_readlane((int)wb, (int)mi);
                    }
                    m = uni(m);
                    w = uni(w);
                    wave_copy<RT>(sh, o, pos + m, g0, w >> 16, w & 0xFFFFu, last_marker_end);
                }
            }
            pos += run;
            if (eob) {
                r = R_EOB;
                break;
            }
        }
        if (is_wide<RT>(o) && o.seg0 + pos >= WIN && o.seg0 + pos - last_marker_end >= WIN) {
            r = R_FLIP;
            break;
        }
        if (pos >= 0x40000000u) {
            // rebase the segment so positions stay 32-bit
            o.p = o.p + (is_wide<RT>(o) ? 2ull : 1ull) * pos;
                        o.seg0 += pos;
            o.cap -= pos;
            g0 += pos;
            cap = (uint32_t)(o.cap < 0x7FFFFFFFull ? o.cap : 0x7FFFFFFFull);
            pos = 0;
        }
    }
    o.pos = pos;
    return r;
}

// Stored block body: LEN bytes from the input at byte `src_byte`.  The wave
// moves 1 KiB per step: every lane issues its aligned dword loads for 4
// four-byte groups before the one wait, realigns them with v_alignbit and
// writes the output (one dword store per group when the output is dword
// aligned and the entries are bytes).  Only the block's last RING entries
// go to the ring; earlier ones would be overwritten there.
template <typename RT>
__device__ __attribute__((always_inline)) void stored_copy(const uint8_t *in_, uint64_t n, uint64_t src_byte, RT *ring,
                                                           const OutSeg &o, uint32_t len) {
    constexpr uint32_t G = 4;  // groups per lane per step (more would raise the kernel's VGPR peak)
    const uint32_t lane = lane_id();
    gptr<const uint32_t> inw = (gptr<const uint32_t>)to_global(in_);
    const uint64_t gpos = o.seg0 + o.pos;
    const uint32_t sh = (uint32_t)(src_byte & 3) * 8;
    const uint64_t w0 = src_byte >> 2;
    const uint64_t nwords = n >> 2;  // whole words inside the buffer
    const uint32_t ring_from = len > RING<RT> ? len - RING<RT> : 0;
    const bool dst_aligned = !is_wide<RT>(o) && (((uint64_t)o.p + o.pos) & 3) == 0;
    for (uint32_t k = 0; k < len; k += G * WAVE * 4) {
        uint32_t lo[G], hi[G];
        const uint64_t qk = (uint64_t)(k >> 2);
        if (w0 + qk + G * WAVE + 1 < nwords) {
#pragma unroll
            for (uint32_t g = 0; g < G; g++) {
                const uint64_t wi = w0 + qk + g * WAVE + lane;
                lo[g] = inw[wi];
                hi[g] = inw[wi + 1];
            }
        } else {
#pragma unroll
            for (uint32_t g = 0; g < G; g++) {
                const uint64_t wi = w0 + qk + g * WAVE + lane;
                lo[g] = load_word(in_, n, wi);
                hi[g] = load_word(in_, n, wi + 1);
            }
        }
#pragma unroll
        for (uint32_t g = 0; g < G; g++) {
            const uint32_t t = k + (g * WAVE + lane) * 4;  // first entry of the group
            if (t < len) {
                const uint32_t v = __builtin_amdgcn_alignbit(hi[g], lo[g], sh);
                const uint32_t nb = len - t < 4 ? len - t : 4;
                if (dst_aligned && nb == 4) {
                    *(gptr<uint32_t>)(o.p + o.pos + t) = v;
                } else {
                    for (uint32_t j = 0; j < nb; j++) put<RT>(o, o.pos + t + j, (v >> (8 * j)) & 0xFF);
                }
                if (t + 4 > ring_from) {
                    for (uint32_t j = 0; j < nb; j++)
                        if (t + j >= ring_from) ring[(uint32_t)(gpos + t + j) & RMASK<RT>] = (RT)((v >> (8 * j)) & 0xFF);
                }
            }
        }
    }
    compiler_fence();
}

template <typename RT, uint32_t LW>
__device__ __attribute__((always_inline)) void inflate_piece(InflateShared<RT, LW> &sh, const uint8_t *__restrict__ in,
                                                             uint64_t in_len, const PieceDesc &pd, PieceOut *po,
                                                             uint64_t *prog, float prog_inv_total) {
    const uint32_t lane = lane_id();
    const bool marker = (pd.flags & PF_MARKER) != 0;
    if (marker && sizeof(RT) == 1) {  // host error: marker pieces need the u16-ring kernel
        if (lane == 0) {
            po->status = PS_ERROR;
            po->err = E_NONE;
            po->n16 = po->n8 = 0;
        }
        return;
    }
    gptr<uint8_t> dst = to_global(pd.dst);
    // marker pieces start with u16 entries; everything else writes u8
    OutSeg o;
    o.p = dst;
    o.pos = 0;
    o.seg0 = 0;
    o.wide = marker ? 1u : 0u;
    o.cap = marker ? pd.cap / 2 : pd.cap;
    uint64_t n16 = 0, u8_off = 0;
    uint64_t last_marker_end = 0;
    uint32_t status = PS_NONE, err = E_NONE, nblocks = 0;
    uint64_t end_bit = 0;
    uint64_t t_begin = wall_clock64();
    uint64_t cp_bit = NO_CP, cp_out = 0;
    uint32_t fixed_built = 0;  // the fixed tables are in the LUTs
    Prog pg;
    pg.on = prog != nullptr || pd.weight > 0.0f;
    pg.k = 0;
    if (lane == 0) {
        const uint64_t end = pd.stop_bit < pd.limit_bit ? pd.stop_bit : pd.limit_bit;
        const uint64_t span = end > pd.start_bit ? end - pd.start_bit : 1;
        sh.pg.ctr = prog;
        sh.pg.inv_total = prog_inv_total;
        sh.pg.inv_span = 1.0f / (float)span;
        sh.pg.weight = pd.weight;
        sh.pg.start = sh.pg.last = pd.start_bit;
    }
    if (!prog && pd.weight > 0.0f) prio_by_remaining(pd.weight);  // a new piece of a batch

    // ring = the last RING entries of the window before the piece (piece
    // position a at slot a mod RING)
    if (marker) {
        for (uint32_t q = lane; q < RING<RT>; q += WAVE) sh.ring[q] = (RT)(MARKER_BASE + WIN - RING<RT> + q);
    } else if (pd.flags & PF_WINDOW) {
        for (uint32_t q = WIN - RING<RT> + lane; q < WIN; q += WAVE) sh.ring[q & RMASK<RT>] = to_global(pd.window)[q];
    } else {
        for (uint32_t q = lane; q < RING<RT>; q += WAVE) sh.ring[q] = 0;
    }
    if (lane == 0) {
        sh.far_out16 = (uint64_t)pd.dst;
        sh.far_out8 = (uint64_t)pd.dst;
        sh.far_n16 = marker ? ~0ull : 0ull;
        sh.far_win = (uint64_t)pd.window;
        sh.far_win_kind = marker ? 3u : (pd.flags & PF_WINDOW) ? 1u : 0u;
    }
    __syncthreads();

    // A MIDBLOCK piece reads its block's header first and starts at the sync
    // point of start_bit; stop_at is the exact stop inside the block at
    // stop_hdr (STOPMID), once that block's tables are built.
    bool mid = (pd.flags & PF_MIDBLOCK) != 0;
    uint64_t start_bit = pd.start_bit, stop_at = NO_SYNC;
    BitReader br;
    br.init(in, in_len, mid ? pd.hdr_bit : pd.start_bit);
    for (;;) {
        br.make_uniform();
        br.refill();
        uint64_t hbit = br.pos();
        if (hbit + 3 > pd.limit_bit) {
            status = PS_ERROR;
            err = E_TRUNC;
            end_bit = hbit;
            break;
        }
        uint32_t hdr = br.read(3);
        uint32_t bfinal = hdr & 1, btype = hdr >> 1;
        if (btype == 3) {
            status = PS_ERROR;
            err = E_BTYPE3;
            end_bit = hbit;
            break;
        }
        if (hbit >= pd.stop_bit && !bfinal) {
            status = PS_STOPPED;
            end_bit = hbit;
            break;
        }
        if (cp_bit == NO_CP && hbit >= pd.start_bit) {
            cp_bit = hbit;
            cp_out = o.seg0 + o.pos;
        }
        nblocks++;
        uint32_t r = R_EOB;
        stop_at = NO_SYNC;
        if (btype == 0 && mid) {
            status = PS_ERROR;  // no symbols to synchronise on
            err = E_SYNC;
            end_bit = pd.start_bit;
            break;
        }
        if (btype == 0) {
            uint32_t pad = (uint32_t)((8 - (br.pos() & 7)) & 7);
            br.consume(pad);
            br.refill();
            uint32_t len = br.read(16), nlen = br.read(16);
            uint64_t src_byte = br.pos() >> 3;
            if ((len ^ 0xFFFFu) != nlen) {
                status = PS_ERROR;
                err = E_STORED_LEN;
                end_bit = hbit;
                break;
            }
            if ((src_byte + len) * 8 > pd.limit_bit) {
                status = PS_ERROR;
                err = E_TRUNC;
                end_bit = hbit;
                break;
            }
            if (o.pos + len > o.cap) {
                r = R_OVERFLOW;
            } else {
                stored_copy<RT>(in, in_len, src_byte, sh.ring, o, len);
                o.pos += len;
                if (is_wide<RT>(o) && o.seg0 + o.pos >= WIN && o.seg0 + o.pos - last_marker_end >= WIN) r = R_FLIP;
                if (o.pos >= 0x40000000u) {  // keep segment positions 32-bit
                    o.p = o.p + (is_wide<RT>(o) ? 2ull : 1ull) * o.pos;
                    o.seg0 += o.pos;
                    o.cap -= o.pos;
                    o.pos = 0;
                }
            }
            br.init(in, in_len, (src_byte + len) * 8);
        } else {
            if (btype == 1) {
                if (!fixed_built) build_fixed(sh);
                fixed_built = 1;
            } else {
                fixed_built = 0;
                err = read_dynamic(sh, br);
                if (err) {
                    status = PS_ERROR;
                    end_bit = hbit;
                    break;
                }
            }
            uint64_t P = br.pos();
            const uint64_t first = P;
            if (mid) {
                mid = false;
                P = sync_point(sh, in, in_len, pd.start_bit, first, pd.limit_bit);
                if (P == NO_SYNC) {
                    status = PS_ERROR;
                    err = E_SYNC;
                    end_bit = pd.start_bit;
                    break;
                }
                start_bit = P;
            }
            if ((pd.flags & PF_STOPMID) && hbit == pd.stop_hdr) {
                stop_at = sync_point(sh, in, in_len, pd.stop_bit, first, pd.limit_bit);
                if (stop_at != NO_SYNC && stop_at <= P) stop_at = NO_SYNC;
            }
            r = decode_symbols<RT>(sh, in, in_len, P, o, pd.limit_bit, last_marker_end, err, stop_at, pg);
            br.init(in, in_len, P);
        }
        if (r == R_STOPMID) {
            status = PS_STOPPED;
            end_bit = br.pos();
            break;
        }
        if (r == R_FLIP) {
            // the ring already holds the last entries, all literal: continue
            // with u8 entries after the u16 ones (slab byte u8_off on)
            n16 = o.seg0 + o.pos;
            u8_off = (2 * n16 + 255) & ~255ull;
            if (u8_off >= pd.cap) {
                r = R_OVERFLOW;
            } else {
                o.p = dst + u8_off;
                o.seg0 = n16;
                o.cap = pd.cap - u8_off;
                o.pos = 0;
                o.wide = 0;
                if (lane == 0) {
                    sh.far_n16 = n16;
                    sh.far_out8 = (uint64_t)pd.dst + u8_off;
                }
                __syncthreads();
                // a flip inside a block resumes that block's symbols
                if (btype != 0) {
                    uint64_t P = br.pos();
                    r = decode_symbols<RT>(sh, in, in_len, P, o, pd.limit_bit, last_marker_end, err, stop_at, pg);
                    br.init(in, in_len, P);
                } else {
                    r = R_EOB;
                }
            }
        }
        if (r == R_STOPMID) {
            status = PS_STOPPED;
            end_bit = br.pos();
            break;
        }
        if (r == R_OVERFLOW) {
            status = PS_OVERFLOW;
            end_bit = br.pos();
            break;
        }
        if (r == R_ERROR) {
            status = PS_ERROR;
            end_bit = br.pos();
            break;
        }
        if (br.pos() > pd.limit_bit) {
            status = PS_ERROR;
            err = E_TRUNC;
            end_bit = br.pos();
            break;
        }
        if (bfinal) {
            status = PS_FINAL;
            end_bit = br.pos();
            break;
        }
    }
    __syncthreads();
    if (lane == 0) {
        const bool flipped = marker && o.wide == 0;
        po->start_bit = start_bit;
        po->end_bit = end_bit;
        const uint64_t total = o.seg0 + o.pos;
        po->n16 = marker ? (flipped ? n16 : total) : 0;
        po->n8 = marker ? (flipped ? total - n16 : 0) : total;
        po->u8_off = flipped ? u8_off : 0;
        po->bits_done = br.pos() > start_bit ? br.pos() - start_bit : 0;
        po->t_begin = t_begin;
        po->t_end = wall_clock64();
        po->status = status;
        po->err = err;
        po->nblocks = nblocks;
        po->mode = marker ? (flipped ? 2u : 1u) : 0u;
        po->cp_bit = cp_bit;
        po->cp_out = cp_out;
    }
}

// Persistent inflate kernel: every workgroup (one wavefront) pulls piece
// indices from a device queue until it is drained; `order` lists the pieces
// longest-first.  RT = ring entry type: uint16_t for MARKER pieces, uint8_t
// for all others; LW the litlen LUT root width.  n_dev (when not null): the
// piece count is read from device memory (a launch inside a graph whose list
// a kernel before it wrote), n is its capacity.  prog (when not null): the
// launch-wide progress counter (zeroed before the launch) and 1 / the bits of
// all its pieces, for the issue-priority balance (progress_report).
template <typename RT, uint32_t LW = 9>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_inflate(const uint8_t *__restrict__ in, uint64_t in_len,
                                                const PieceDesc *__restrict__ pieces,
                                                const uint32_t *__restrict__ order, uint32_t n, uint32_t *queue,
                                                PieceOut *out, const uint32_t *__restrict__ n_dev,
                                                uint64_t *prog, float prog_inv_total) {
    __shared__ InflateShared<RT, LW> sh;
    if (n_dev) {
        const uint32_t m = uni(*n_dev);
        n = m < n ? m : n;
    }
    for (;;) {
        if (lane_id() == 0) sh.qi = atomicAdd(queue, 1u);
        __syncthreads();
        uint32_t qi = uni(sh.qi);
        __syncthreads();
        if (qi >= n) break;
        uint32_t pi = order[qi];
        PieceDesc pd = pieces[pi];
        inflate_piece<RT, LW>(sh, in, in_len, pd, out + pi, prog, prog_inv_total);
        __syncthreads();
    }
}

}  // namespace gz
