// k_resolve.h — marker resolution and window propagation.
//
// A marker piece's output is u16 entries (< 256 literal, >= 32768 marker =
// index into the 32 KiB window that preceded the piece) followed, after the
// flip, by clean u8 bytes.  Resolution is MapMarkers
// (replace_markers.rs:47,129; apply_window.rs:22):
//     out[i] = v < 256 ? v : window[v - 32768]
// with window = the 32768 final output bytes before the piece (zero-padded
// at the member start, WindowMap{0: zeros}).
//
// A piece that flipped to u8 entries has final bytes as its last 32768
// entries.  When those are all u8 entries and no later window reaches back
// into them, k_resolve_body resolves that piece whole and its successor
// reads its window from the piece's slab.  Every other marker piece belongs
// to a chain: k_chain walks each chain in order, keeping the rolling 32 KiB
// window in LDS, and writes each piece's last 32 KiB (the end window its
// successor needs: get_last_window, chunk_data.rs:1295); a chain starts anew
// at a flipped piece, and chains are independent workgroups.  k_resolve_body
// then resolves every entry in parallel, reading each piece's window from
// the output (final by then) or from that flipped predecessor's slab; it
// writes the end windows k_chain wrote once more, with the same bytes (an
// entry depends only on the window before its piece), so that their CRC is
// folded in the same pass.
#pragma once
#include "gz_common.h"
#include "k_crc.h"

namespace gz {

struct ResolvePiece {
    const uint8_t *slab;
    uint8_t *mout;      // member output base
    uint64_t out_pos;   // piece offset in the member output
    uint64_t n16, n8, u8_off;
    uint64_t tail;      // member output bytes after the piece
    uint64_t body;      // entries [0, body) are k_resolve_body's (all of them)
    // the byte after the 32 KiB window before the piece: in the member
    // output (win_out, positions before the member start read 0), or the
    // end of the predecessor's u8 entries in its slab when that piece
    // flipped (its last 32768 entries are final bytes already)
    const uint8_t *win_end;
    uint32_t win_out, pad;
};

// Body tiles: entries [RESOLVE_TILE k, RESOLVE_TILE (k+1)) of a piece's
// first `body` entries; tile_first[i] is the index of piece i's first tile
// (a prefix sum, tile_first[npieces] = all tiles).
constexpr uint64_t RESOLVE_TILE = 64 * 1024;

__device__ __forceinline__ uint8_t piece_entry(const ResolvePiece &r, uint64_t t, const uint8_t *win) {
    if (t < r.n16) {
        uint32_t e = ((gptr<const uint16_t>)to_global(r.slab))[t];
        return e < 256 ? (uint8_t)e : win[e - MARKER_BASE];
    }
    return to_global(r.slab)[r.u8_off + (t - r.n16)];
}

__global__ void __launch_bounds__(1024) k_chain(const ResolvePiece *__restrict__ rp, const uint32_t *__restrict__ chains,
                                                uint32_t nchains) {
    __shared__ uint8_t win[2][WIN];
    for (uint32_t c = blockIdx.x; c < nchains; c += gridDim.x) {
        const uint32_t p0 = chains[2 * c], p1 = chains[2 * c + 1];
        {
            // a chain that starts at a flipped piece (u8 entries present,
            // >= 32768 entries) never reads the window before it: leave it
            // unloaded, since that output may still be being written by
            // another chain
            const ResolvePiece f = rp[p0];
            if (!(f.u8_off != 0 && f.n16 + f.n8 >= WIN))
                for (uint32_t j = threadIdx.x; j < WIN; j += blockDim.x) {
                    const int64_t q = (int64_t)j - (int64_t)WIN;
                    win[0][j] = f.win_out && (int64_t)f.out_pos + q < 0 ? 0 : to_global(f.win_end)[q];
                }
        }
        __syncthreads();
        uint32_t cur = 0;
        for (uint32_t i = p0; i < p1; i++) {
            const ResolvePiece r = rp[i];
            const uint64_t L = r.n16 + r.n8;
            for (uint32_t j = threadIdx.x; j < WIN; j += blockDim.x) {
                int64_t t = (int64_t)L - (int64_t)WIN + j;
                uint8_t v;
                if (t < 0) {
                    v = win[cur][WIN + t];
                } else {
                    v = piece_entry(r, (uint64_t)t, win[cur]);
                    to_global(r.mout)[r.out_pos + t] = v;
                }
                win[cur ^ 1][j] = v;
            }
            __syncthreads();
            cur ^= 1;
        }
        __syncthreads();
    }
}

// ---- the window chain as a parallel prefix (long chains) -----------------
// k_chain walks a chain piece by piece; a chain of thousands of pieces (a
// member of a few hundred MiB, whose pieces are too short to flip) then takes
// tens of milliseconds.  Instead: slot s holds the end window of piece
// rp[slot_rp[s]], the slot_pos[s]-th piece of its chain (a chain's slots are
// consecutive), as WIN u16 entries: < 256 a byte, MARKER_BASE + m byte m of
// the window before that piece (scan_raw: the piece's own last WIN entries,
// and for a piece shorter than WIN the window bytes before it as markers); a
// chain's first piece is resolved at once from its own window.  Step d
// replaces every marker of slot s (slot_pos[s] >= d) by entry m of slot
// s - d, the window its markers refer to, then the markers left by entry m
// of slot s - 2d, and so on up to slot s - (CHAIN_RADIX - 1) d (each window
// in the state before the step); steps go d = 1, 16, 256, ...  After step d,
// slots at positions < CHAIN_RADIX d hold bytes only, so
// max(1, ceil(log16(longest chain))) steps resolve every slot (fewer passes
// over the slots; most markers resolve at the first window back, so the
// deeper look-back of a larger radix costs little: a 256 MiB text member's
// chain of 3400 pieces in 3 steps took 0.90 ms of resolve where radix 8's 4
// steps took 0.99 ms).  The first step reads the slots'
// entries from the slabs (scan_raw) rather than from a copy of them, the
// last one writes the resolved bytes to the member output rather than to
// the slot buffer; steps between are double-buffered, each reading one
// buffer and writing the other.

constexpr uint32_t CHAIN_RADIX = 16;

// Entry j (< WIN) of the end window of piece r, before any step.
__device__ __forceinline__ uint32_t scan_raw(const ResolvePiece &r, bool chain_first, uint32_t j) {
    const uint64_t L = r.n16 + r.n8;
    const int64_t t = (int64_t)L - (int64_t)WIN + (int64_t)j;
    uint32_t e;
    if (t < 0) e = MARKER_BASE + (uint32_t)(L + j);  // window byte L + j (position t)
    else if ((uint64_t)t < r.n16) e = ((gptr<const uint16_t>)to_global(r.slab))[t];
    else e = to_global(r.slab)[r.u8_off + ((uint64_t)t - r.n16)];
    if (chain_first && e >= MARKER_BASE) {
        const int64_t q = (int64_t)(e - MARKER_BASE) - (int64_t)WIN;
        e = r.win_out && (int64_t)r.out_pos + q < 0 ? 0u : (uint32_t)to_global(r.win_end)[q];
    }
    return e;
}

template <bool FIRST, bool LAST>
__global__ void __launch_bounds__(256) k_chain_scan_step(const ResolvePiece *__restrict__ rp,
                                                         const uint32_t *__restrict__ slot_rp,
                                                         const uint32_t *__restrict__ slot_pos, uint32_t nslots,
                                                         uint32_t step, const uint16_t *__restrict__ Win,
                                                         uint16_t *__restrict__ Wout) {
    typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
    // FIRST: the pieces of slots s, s - 1, ..., s - (CHAIN_RADIX - 1)
    __shared__ ResolvePiece prs[CHAIN_RADIX];
    for (uint32_t s = blockIdx.x; s < nslots; s += gridDim.x) {
        // up to CHAIN_RADIX - 1 windows back: s - step, s - 2 step, ...
        const uint32_t pos = slot_pos[s];
        const uint32_t back = pos >= (CHAIN_RADIX - 1) * step ? CHAIN_RADIX - 1 : pos / step;
        if (FIRST || LAST) {
            __syncthreads();  // the previous slot's readers of prs are done
            if (threadIdx.x <= (FIRST ? back : 0)) prs[threadIdx.x] = rp[slot_rp[s - threadIdx.x]];
            __syncthreads();
        }
        gptr<const u16x8> src = (gptr<const u16x8>)(to_global(Win) + (uint64_t)s * WIN);
        gptr<u16x8> dst = (gptr<u16x8>)(to_global(Wout) + (uint64_t)s * WIN);
        for (uint32_t v = threadIdx.x; v < WIN / 8; v += blockDim.x) {
            u16x8 x;
            if (FIRST) {
                const ResolvePiece &r = prs[0];
                const int64_t t0 = (int64_t)(r.n16 + r.n8) - (int64_t)WIN + 8 * (int64_t)v;
                if (t0 >= 0 && (pos != 0 || (uint64_t)t0 >= r.n16)) {
                    // eight u16 entries, or eight u8 entries, or the seam:
                    // aligned dword loads from the slab, realigned
                    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                    u32x4 e;
                    if ((uint64_t)t0 + 8 <= r.n16) {
                        const uint64_t so = 2 * (uint64_t)t0;
                        gptr<const uint32_t> sw = (gptr<const uint32_t>)(to_global(r.slab) + (so & ~3ull));
                        const uint32_t sh = (uint32_t)(so & 3) * 8;
                        const uint32_t w0 = sw[0], w1 = sw[1], w2 = sw[2], w3 = sw[3], w4 = sh ? sw[4] : 0;
                        e.x = __builtin_amdgcn_alignbit(w1, w0, sh);
                        e.y = __builtin_amdgcn_alignbit(w2, w1, sh);
                        e.z = __builtin_amdgcn_alignbit(w3, w2, sh);
                        e.w = __builtin_amdgcn_alignbit(w4, w3, sh);
                        x = __builtin_bit_cast(u16x8, e);
                    } else if ((uint64_t)t0 >= r.n16) {
                        const uint64_t so = r.u8_off + ((uint64_t)t0 - r.n16);
                        gptr<const uint32_t> sw = (gptr<const uint32_t>)(to_global(r.slab) + (so & ~3ull));
                        const uint32_t sh = (uint32_t)(so & 3) * 8;
                        const uint32_t w0 = sw[0], w1 = sw[1], w2 = sh ? sw[2] : 0;
                        const uint32_t lo = __builtin_amdgcn_alignbit(w1, w0, sh);
                        const uint32_t hi = __builtin_amdgcn_alignbit(w2, w1, sh);
                        // bytes to u16 entries
                        e.x = __builtin_amdgcn_perm(0u, lo, 0x0c010c00u);
                        e.y = __builtin_amdgcn_perm(0u, lo, 0x0c030c02u);
                        e.z = __builtin_amdgcn_perm(0u, hi, 0x0c010c00u);
                        e.w = __builtin_amdgcn_perm(0u, hi, 0x0c030c02u);
                        x = __builtin_bit_cast(u16x8, e);
                    } else {
#pragma unroll
                        for (int k = 0; k < 8; k++) x[k] = (uint16_t)scan_raw(r, false, 8 * v + k);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < 8; k++) x[k] = (uint16_t)scan_raw(r, pos == 0, 8 * v + k);
                }
            } else {
                x = src[v];
            }
            for (uint32_t b = 1; b <= back; b++) {
                bool more = false;
                if (FIRST) {
                    // step 1: slot s - b's entries straight from its slab
                    const bool cf = b == pos;
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        if (x[k] >= MARKER_BASE) x[k] = (uint16_t)scan_raw(prs[b], cf, x[k] - MARKER_BASE);
                        more |= x[k] >= MARKER_BASE;
                    }
                } else {
                    gptr<const uint16_t> prev = to_global(Win) + (uint64_t)(s - b * step) * WIN;
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        if (x[k] >= MARKER_BASE) x[k] = prev[x[k] - MARKER_BASE];
                        more |= x[k] >= MARKER_BASE;
                    }
                }
                if (!__any(more)) break;  // wave-uniform exit
            }
            if (LAST) {
                // every entry is a byte now: the piece's entries (the last
                // WIN, or all of a shorter piece) go to the member output
                const ResolvePiece &r = prs[0];
                const int64_t t0 = (int64_t)(r.n16 + r.n8) - (int64_t)WIN + 8 * (int64_t)v;
                gptr<uint8_t> o = to_global(r.mout) + (int64_t)r.out_pos + t0;
                if (t0 >= 0 && ((uintptr_t)o & 7) == 0) {
                    uint64_t q = 0;
#pragma unroll
                    for (int k = 0; k < 8; k++) q |= (uint64_t)(x[k] & 0xFF) << (8 * k);
                    *(gptr<uint64_t>)o = q;
                } else {
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        if (t0 + k >= 0) o[k] = (uint8_t)x[k];
                }
            } else {
                dst[v] = x;
            }
        }
    }
}

// Launches the steps over nslots slots (two slot buffers of nslots * WIN
// u16 entries each); returns the number of launches.
inline uint32_t launch_chain_scan(hipStream_t stream, dim3 grid, const ResolvePiece *rp, const uint32_t *slot_rp,
                                  const uint32_t *slot_pos, uint32_t nslots, uint32_t max_chain, uint16_t *const wb[2]) {
    uint32_t n = 0, cur = 0;
    for (uint32_t d = 1; n == 0 || d < max_chain; d *= CHAIN_RADIX, n++, cur ^= 1) {
        const bool first = n == 0, last = (uint64_t)d * CHAIN_RADIX >= max_chain;
        auto k = first ? (last ? k_chain_scan_step<true, true> : k_chain_scan_step<true, false>)
                       : (last ? k_chain_scan_step<false, true> : k_chain_scan_step<false, false>);
        hipLaunchKernelGGL(k, grid, dim3(256), 0, stream, rp, slot_rp, slot_pos, nslots, d,
                           (const uint16_t *)wb[cur], wb[cur ^ 1]);
    }
    return n;
}

// Value of entry t of a piece for the body pass: window bytes come from the
// 32 KiB before the piece (win_end; zeros before the member start).
__device__ __forceinline__ uint8_t body_entry(const ResolvePiece &r, uint64_t t) {
    if (t < r.n16) {
        const uint32_t e = ((gptr<const uint16_t>)to_global(r.slab))[t];
        if (e < 256) return (uint8_t)e;
        const int64_t q = (int64_t)(e - MARKER_BASE) - (int64_t)WIN;
        return r.win_out && (int64_t)r.out_pos + q < 0 ? 0 : to_global(r.win_end)[q];
    }
    return to_global(r.slab)[r.u8_off + (t - r.n16)];
}

__device__ __forceinline__ uint32_t resolve16(uint32_t e, gptr<const uint8_t> win, bool near_start, int64_t out_pos) {
    if (e < 256) return e;
    const int64_t q = (int64_t)(e - MARKER_BASE) - (int64_t)WIN;
    return (near_start && out_pos + q < 0) ? 0u : (uint32_t)win[q];
}

// Body pass: every entry of a marker piece (the last 32768 of a chained
// piece again: k_chain wrote them first, the same bytes).  One thread per 16-byte-aligned block of the output, written
// with one 16-byte store when the block lies inside the tile:
//   * all u8 entries: five aligned dword loads from the slab, realigned with
//     v_alignbyte (the slab's u8 part is 256-byte aligned, the output is not);
//   * all u16 entries: nine aligned dword loads, markers resolved from the
//     window just before the piece;
//   * tile edges and the u16/u8 seam: entry by entry.
// The tile's CRC32 comes out of the same pass (the bytes are not read
// again): over the tile's 16-byte blocks [A0, A1) with the bytes outside
// the tile taken as zero, block i contributes crc(block) * x^(8 (A1 - end of
// block)).  A thread's blocks are 4096 bytes apart, so it folds them in
// order with acc = acc * x^(8 * 4096) ^ crc(block) and shifts once to A1;
// the threads' values XOR together, the trailing zeros (A1 - d1) are divided
// out and the result is shifted over the member bytes after the tile, as
// k_crc_segments does for its segments.
__global__ void __launch_bounds__(256) k_resolve_body(const ResolvePiece *__restrict__ rp,
                                                      const uint32_t *__restrict__ tile_first, uint32_t npieces,
                                                      const uint32_t *__restrict__ crc_tables,
                                                      uint32_t *__restrict__ tile_crc, uint32_t t_begin,
                                                      uint32_t t_end) {
    const uint32_t ntiles = tile_first[npieces] < t_end ? tile_first[npieces] : t_end;
    __shared__ uint32_t t8[8][256];
    __shared__ uint32_t mx[4][256];
    __shared__ uint32_t red;
    for (uint32_t i = threadIdx.x; i < 8 * 256; i += blockDim.x) t8[i >> 8][i & 255] = crc_tables[i];
    for (uint32_t i = threadIdx.x; i < 4 * 256; i += blockDim.x) mx[i >> 8][i & 255] = crc_tables[8 * 256 + i];
    __syncthreads();
    for (uint32_t b = t_begin + blockIdx.x; b < ntiles; b += gridDim.x) {
        if (threadIdx.x == 0) red = 0;
        __syncthreads();
        uint32_t acc = 0;
        int64_t my_end = -1;  // end of this thread's last block
        // the tile's piece: the last one whose first tile is at or before b
        uint32_t lo = 0, hi = npieces;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (tile_first[mid] <= b) lo = mid;
            else hi = mid;
        }
        const ResolvePiece r = rp[lo];
        const uint64_t body = r.body;
        const uint64_t tl_t0 = (uint64_t)(b - tile_first[lo]) * RESOLVE_TILE;
        const uint64_t tl_t1 = tl_t0 + RESOLVE_TILE < body ? tl_t0 + RESOLVE_TILE : body;
        const uint64_t tl_tail = r.tail + (r.n16 + r.n8 - tl_t1);
        gptr<uint8_t> mout = to_global(r.mout);
        gptr<const uint8_t> win = to_global(r.win_end);  // window = win[-WIN .. 0)
        const bool near_start = r.win_out && r.out_pos < WIN;
        const int64_t d0 = (int64_t)(r.out_pos + tl_t0), d1 = (int64_t)(r.out_pos + tl_t1);  // output range
        // blocks are 16-byte aligned in the address space (the member's
        // output base need not be aligned); offsets relative to mout
        const int64_t ba = (int64_t)(uintptr_t)r.mout;
        const int64_t first = ((ba + d0) & ~(int64_t)15) - ba;
        for (int64_t blk = first + 16 * (int64_t)threadIdx.x; blk < d1; blk += 16 * (int64_t)blockDim.x) {
            const bool full = blk >= d0 && blk + 16 <= d1;
            const uint64_t t = (uint64_t)(blk - (int64_t)r.out_pos);  // entry of the block's first byte (full blocks)
            if (full && t >= r.n16) {
                // u8 entries [t, t + 16) at slab offset u8_off + t - n16
                const uint64_t so = r.u8_off + (t - r.n16);
                gptr<const uint32_t> sw = (gptr<const uint32_t>)(to_global(r.slab) + (so & ~3ull));
                const uint32_t sh = (uint32_t)(so & 3) * 8;
                uint32_t x0 = sw[0], x1 = sw[1], x2 = sw[2], x3 = sw[3], x4 = sh ? sw[4] : 0;
                typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                u32x4 o;
                o.x = __builtin_amdgcn_alignbit(x1, x0, sh);
                o.y = __builtin_amdgcn_alignbit(x2, x1, sh);
                o.z = __builtin_amdgcn_alignbit(x3, x2, sh);
                o.w = __builtin_amdgcn_alignbit(x4, x3, sh);
                *(gptr<u32x4>)(mout + blk) = o;
                acc = mul_x4k(mx, acc) ^ crc16b(t8, o.x, o.y, o.z, o.w);
            } else if (full && t + 16 <= r.n16) {
                // u16 entries [t, t + 16): 32 bytes at slab offset 2t
                const uint64_t so = 2 * t;
                gptr<const uint32_t> sw = (gptr<const uint32_t>)(to_global(r.slab) + (so & ~3ull));
                const uint32_t sh = (uint32_t)(so & 3) * 8;
                uint32_t x[9];
#pragma unroll
                for (int k = 0; k < 8; k++) x[k] = sw[k];
                x[8] = sh ? sw[8] : 0;
                uint32_t e[8];
                uint32_t mk = 0;  // bit 15 of an entry: a marker (literals are below 256)
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    e[k] = __builtin_amdgcn_alignbit(x[k + 1], x[k], sh);
                    mk |= e[k];
                }
                uint32_t ob[4];
                if (!ballot((mk & 0x80008000u) != 0)) {
                    // no marker in the wave's blocks: the low bytes, two
                    // dwords of entries into one of bytes
#pragma unroll
                    for (int k = 0; k < 4; k++) ob[k] = __builtin_amdgcn_perm(e[2 * k + 1], e[2 * k], 0x06040200u);
                } else if (!near_start) {
                    // markers: every entry reads the window byte its value
                    // names, literals a fixed byte inside it (no lane-masked
                    // branches): window byte m of marker 32768 + m is at
                    // win[m - 32768], i.e. at wbase + value
                    gptr<const uint8_t> wbase = win - 2 * (int64_t)WIN;
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        uint32_t o = 0;
#pragma unroll
                        for (int h = 0; h < 4; h++) {
                            const uint32_t v = (e[2 * k + (h >> 1)] >> (16 * (h & 1))) & 0xFFFFu;
                            const uint32_t wv = wbase[(uint64_t)max(v, MARKER_BASE)];
                            o |= (v < 256 ? v : wv) << (8 * h);
                        }
                        ob[k] = o;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const uint32_t lo = e[2 * k], hi = e[2 * k + 1];
                        const uint32_t b0 = resolve16(lo & 0xFFFF, win, near_start, (int64_t)r.out_pos);
                        const uint32_t b1 = resolve16(lo >> 16, win, near_start, (int64_t)r.out_pos);
                        const uint32_t b2 = resolve16(hi & 0xFFFF, win, near_start, (int64_t)r.out_pos);
                        const uint32_t b3 = resolve16(hi >> 16, win, near_start, (int64_t)r.out_pos);
                        ob[k] = b0 | b1 << 8 | b2 << 16 | b3 << 24;
                    }
                }
                typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                u32x4 o;
                o.x = ob[0];
                o.y = ob[1];
                o.z = ob[2];
                o.w = ob[3];
                *(gptr<u32x4>)(mout + blk) = o;
                acc = mul_x4k(mx, acc) ^ crc16b(t8, o.x, o.y, o.z, o.w);
            } else {
                uint32_t w[4] = {0, 0, 0, 0};
                for (int64_t k = 0; k < 16; k++) {
                    const int64_t d = blk + k;
                    if (d >= d0 && d < d1) {
                        const uint8_t v = body_entry(r, (uint64_t)(d - (int64_t)r.out_pos));
                        mout[d] = v;
                        w[k >> 2] |= (uint32_t)v << (8 * (k & 3));
                    }
                }
                acc = mul_x4k(mx, acc) ^ crc16b(t8, w[0], w[1], w[2], w[3]);
            }
            my_end = blk + 16;
        }
        // the tile's blocks end at A1 (16-byte aligned past d1)
        const int64_t A1 = first + ((d1 - first + 15) & ~(int64_t)15);
        if (my_end >= 0) acc = crc_shift(acc, (uint64_t)(A1 - my_end));
        if (acc) atomicXor(&red, acc);
        __syncthreads();
        if (threadIdx.x == 0) tile_crc[b] = crc_shift(crc_unshift(red, (uint32_t)(A1 - d1)), tl_tail);
        __syncthreads();
    }
}

}  // namespace gz
