// k_inflate.h — speculative chunk inflater: one wavefront per piece.
//
// A "piece" is a bit range of one DEFLATE stream that starts at a block
// header (known from the stream start or from the block finder) and ends at
// the first non-final block header at or after its stop bit, or at the end of
// the final block (stop rule of chunk_decode.rs:2029-2040).
//
// Output modes
//   DIRECT : window is all zeros (stream start, WindowMap{0: zeros},
//            chunk_fetcher.rs:694-696); bytes go straight to the member's
//            final output.  Back-references before the start read zero.
//   MARKER : window unknown; u16 entries into the piece's slab: literals
//            < 256, window markers 32768 + index-from-oldest
//            (replace_markers.rs:6-23; init_marker_zone,
//            marker_inflate.rs:255-268).  The slab starts with a 32768-entry
//            u16 preamble holding the marker values themselves, so a copy
//            that reaches before the piece start copies markers for free.
//   CLEAN  : after the last 32 KiB of a MARKER piece hold no marker
//            (WidthRing::should_flip, width_ring.rs:310), the tail is
//            narrowed into a u8 preamble and decoding continues in u8.
//
// Slab layout (bytes): [0, 64 KiB) u16 preamble | u16 marker entries |
// (256-aligned) u8 region: 32 KiB u8 preamble | clean bytes.
#pragma once
#include "gz_common.h"

namespace gz {

// PF_DIRECT: all-zero window, no preamble.  PF_WINDOW: clean u8 decode whose
// 32 KiB window the caller placed at dst[-32768, 0).  PF_MARKER: window absent.
enum : uint32_t { PF_DIRECT = 1, PF_MARKER = 2, PF_WINDOW = 4 };

struct PieceDesc {
    uint64_t start_bit;   // first block header to decode
    uint64_t stop_bit;    // stop before a non-final block header at >= stop_bit
    uint64_t limit_bit;   // end of the member's readable deflate bits
    uint8_t *dst;         // DIRECT: output bytes; MARKER: slab base
    uint64_t cap;         // DIRECT: output capacity (bytes); MARKER: slab bytes
    uint32_t flags;
    uint32_t member;
};

struct PieceOut {
    uint64_t end_bit;
    uint64_t n16;       // u16 marker-mode entries
    uint64_t n8;        // u8 entries (DIRECT, or CLEAN after the flip)
    uint64_t u8_off;    // slab byte offset of the u8 region (its 32 KiB preamble first)
    uint64_t bits_done; // bits consumed when the piece stopped (sizing on overflow)
    uint32_t status;
    uint32_t err;
    uint32_t nblocks;
    uint32_t pad;
};

struct __align__(16) InflateShared {
    uint32_t ll[1 << LL_ROOT];
    uint32_t dl[1 << D_ROOT];
    uint16_t ll_sorted[288];
    uint16_t d_sorted[32];
    uint16_t ll_count[16];
    uint16_t d_count[16];
    uint16_t offs[16];
    uint16_t firstc[16];
    uint8_t lens[320 + 2 * WAVE];
    uint8_t pre_lens[32];
    uint32_t qi;
    int32_t fixed;
};

// ---- canonical table construction (whole wave) ----------------------------
// mode 0: litlen, 1: distance, 2: precode (entry = len | sym << 16).
// Returns 0, or -1 when the lengths over-subscribe the code space
// (build_precode_table, marker_inflate.rs:4515-4540).  Incomplete codes are
// accepted; their unused codewords fail when decoded.
template <int MODE>
__device__ int build_table(InflateShared &sh, const uint8_t *lens, int n, uint32_t *lut, int R, uint16_t *count,
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
    return 0;
}

__device__ void build_fixed(InflateShared &sh) {
    const uint32_t lane = lane_id();
    for (uint32_t s = lane; s < 288; s += WAVE) sh.lens[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8;
    __syncthreads();
    build_table<0>(sh, sh.lens, 288, sh.ll, LL_ROOT, sh.ll_count, sh.ll_sorted);
    for (uint32_t s = lane; s < 32; s += WAVE) sh.lens[s] = 5;
    __syncthreads();
    build_table<1>(sh, sh.lens, 32, sh.dl, D_ROOT, sh.d_count, sh.d_sorted);
}

// read_dynamic_huffman_coding (marker_inflate.rs:1206-1245) and
// read_literal_and_distance_code_lengths (:4447-4510).  Returns E_NONE or an
// error code.
__device__ uint32_t read_dynamic(InflateShared &sh, BitReader &br) {
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
    // precode LUT lives in the litlen LUT storage until the litlen table is built
    if (build_table<2>(sh, sh.pre_lens, 19, sh.ll, PRE_ROOT, sh.d_count, sh.d_sorted)) return E_PRECODE;
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
    if (sh.lens[256] == 0) return E_EOB_ZERO;
    if (build_table<0>(sh, sh.lens, (int)hlit, sh.ll, LL_ROOT, sh.ll_count, sh.ll_sorted)) return E_LL_TABLE;
    if (build_table<1>(sh, sh.lens + hlit, (int)hdist, sh.dl, D_ROOT, sh.d_count, sh.d_sorted)) return E_D_TABLE;
    return E_NONE;
}

// ---- LZ77 copy by the whole wave -----------------------------------------
// Entry pos+i takes the value at pos - dist + (i mod dist): every source lies
// before pos, so the lanes of one copy never depend on each other.  Copies
// that read what an earlier symbol wrote rely on the wave's vector memory
// operations completing in program order.
template <typename T, bool MARKER, bool ZW>
__device__ __forceinline__ void wave_copy(T *out, uint64_t pos, uint32_t len, uint32_t dist,
                                          uint64_t &last_marker_end) {
    const uint32_t lane = lane_id();
    const bool overlap = dist < len;
    for (uint32_t k = 0; k < len; k += WAVE) {
        uint32_t i = k + lane;
        bool act = i < len;
        uint32_t si = overlap ? (i % dist) : i;
        int64_t src = (int64_t)pos + (int64_t)si - (int64_t)dist;
        T v = 0;
        if (act) {
            if (!ZW || src >= 0) v = out[src];
            out[pos + i] = v;
        }
        if (MARKER) {
            uint64_t m = ballot(act && v >= 256);
            if (m) last_marker_end = pos + k + 64 - (uint64_t)__clzll(m);
        }
    }
    compiler_fence();
}

enum : uint32_t { R_EOB = 0, R_FLIP = 1, R_OVERFLOW = 2, R_ERROR = 3 };

// Decode Huffman symbols of the current block until end-of-block.
template <typename T, bool MARKER, bool ZW>
__device__ uint32_t decode_symbols(InflateShared &sh, BitReader &br, T *out, uint64_t &pos, uint64_t cap,
                                   uint64_t base, uint64_t limit_bit, uint64_t &last_marker_end, uint32_t &err) {
    const uint32_t lane = lane_id();
    for (;;) {
        if (br.pos() > limit_bit) {
            err = E_TRUNC;
            return R_ERROR;
        }
        br.refill();
        uint32_t e = uni(sh.ll[br.peek(LL_ROOT)]);
        uint32_t L = e & 15;
        if (L == 0) {
            uint32_t sl;
            int s = slow_decode(br.peek(15), sh.ll_count, sh.ll_sorted, sl);
            if (s < 0) {
                err = E_BAD_SYM;
                return R_ERROR;
            }
            e = make_ll_entry((uint32_t)s, sl);
            L = sl;
        }
        uint32_t kind = (e >> 8) & 3;
        if (kind == K_LIT) {
            if (pos >= cap) return R_OVERFLOW;
            br.consume(L);
            if (lane == 0) out[pos] = (T)(e >> 16);
            pos++;
            continue;
        }
        if (kind == K_EOB) {
            br.consume(L);
            compiler_fence();
            return R_EOB;
        }
        if (kind == K_BAD) {
            err = E_BAD_SYM;
            return R_ERROR;
        }
        uint32_t E = (e >> 4) & 15;
        uint32_t len = (e >> 16) + ((uint32_t)(br.buf >> L) & ((1u << E) - 1));
        if (pos + len > cap) return R_OVERFLOW;
        br.consume(L + E);
        br.refill();
        uint32_t d = uni(sh.dl[br.peek(D_ROOT)]);
        uint32_t DL = d & 15;
        if (DL == 0) {
            uint32_t sl;
            int s = slow_decode(br.peek(15), sh.d_count, sh.d_sorted, sl);
            if (s < 0) {
                err = E_BAD_DIST_SYM;
                return R_ERROR;
            }
            d = make_d_entry((uint32_t)s, sl);
            DL = sl;
        }
        if (((d >> 8) & 3) != K_LEN) {
            err = E_BAD_DIST_SYM;
            return R_ERROR;
        }
        uint32_t DE = (d >> 4) & 15;
        uint32_t dist = (d >> 16) + ((uint32_t)(br.buf >> DL) & ((1u << DE) - 1));
        br.consume(DL + DE);
        if ((uint64_t)dist > base + pos + WIN) {
            err = E_DIST_FAR;
            return R_ERROR;
        }
        wave_copy<T, MARKER, ZW>(out, pos, len, dist, last_marker_end);
        pos += len;
        if (MARKER && pos >= WIN && pos - last_marker_end >= WIN) return R_FLIP;
    }
}

// Stored block body: LEN bytes from the input at byte `src_byte`.
template <typename T>
__device__ __forceinline__ void stored_copy(const uint8_t *__restrict__ in, uint64_t src_byte, T *out, uint64_t pos,
                                            uint32_t len) {
    const uint32_t lane = lane_id();
    for (uint32_t k = 0; k < len; k += 8 * WAVE) {
#pragma unroll
        for (uint32_t j = 0; j < 8; j++) {
            uint32_t i = k + j * WAVE + lane;
            if (i < len) out[pos + i] = (T)in[src_byte + i];
        }
    }
    compiler_fence();
}

__device__ void inflate_piece(InflateShared &sh, const uint8_t *__restrict__ in, uint64_t in_len,
                              const PieceDesc &pd, PieceOut *po) {
    const uint32_t lane = lane_id();
    enum { M_DIRECT = 0, M_MARKER = 1, M_CLEAN = 2 };
    int mode = (pd.flags & PF_MARKER) ? M_MARKER : (pd.flags & PF_WINDOW) ? M_CLEAN : M_DIRECT;
    uint64_t clean_base = 0;  // entries before the u8 region (marker part, once flipped)
    uint8_t *dst = pd.dst;
    uint16_t *out16 = nullptr;
    uint8_t *out8 = dst;
    uint64_t pos16 = 0, pos8 = 0, cap16 = 0, cap8 = pd.cap, u8_off = 0;
    uint64_t last_marker_end = 0;
    uint32_t status = PS_NONE, err = E_NONE, nblocks = 0;
    uint64_t end_bit = 0;
    sh.fixed = 0;

    if (mode == M_MARKER) {
        // u16 preamble: entry i (oldest first) holds MARKER_BASE + i
        uint4 *p = reinterpret_cast<uint4 *>(dst);
        for (uint32_t q = lane; q < WIN / 8; q += WAVE) {
            uint32_t b = MARKER_BASE + q * 8;
            uint4 v;
            v.x = b | (b + 1) << 16;
            v.y = (b + 2) | (b + 3) << 16;
            v.z = (b + 4) | (b + 5) << 16;
            v.w = (b + 6) | (b + 7) << 16;
            p[q] = v;
        }
        out16 = reinterpret_cast<uint16_t *>(dst) + WIN;
        cap16 = pd.cap >= 2 * WIN ? (pd.cap - 2 * WIN) / 2 : 0;
        compiler_fence();
    }

    BitReader br;
    br.init(in, in_len, pd.start_bit);
    for (;;) {
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
        nblocks++;
        uint32_t r = R_EOB;
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
            if (mode == M_MARKER) {
                if (pos16 + len > cap16) r = R_OVERFLOW;
                else {
                    stored_copy<uint16_t>(in, src_byte, out16, pos16, len);
                    pos16 += len;
                    if (pos16 >= WIN && pos16 - last_marker_end >= WIN) r = R_FLIP;
                }
            } else {
                uint64_t &p8 = pos8;
                if (p8 + len > cap8) r = R_OVERFLOW;
                else {
                    stored_copy<uint8_t>(in, src_byte, out8, p8, len);
                    p8 += len;
                }
            }
            br.init(in, in_len, (src_byte + len) * 8);
        } else {
            if (btype == 1) {
                if (!sh.fixed) build_fixed(sh);
                sh.fixed = 1;
            } else {
                sh.fixed = 0;
                err = read_dynamic(sh, br);
                if (err) {
                    status = PS_ERROR;
                    end_bit = hbit;
                    break;
                }
            }
            for (;;) {
                if (mode == M_MARKER)
                    r = decode_symbols<uint16_t, true, false>(sh, br, out16, pos16, cap16, 0, pd.limit_bit,
                                                               last_marker_end, err);
                else if (mode == M_CLEAN)
                    r = decode_symbols<uint8_t, false, false>(sh, br, out8, pos8, cap8, clean_base, pd.limit_bit,
                                                              last_marker_end, err);
                else
                    r = decode_symbols<uint8_t, false, true>(sh, br, out8, pos8, cap8, 0, pd.limit_bit,
                                                             last_marker_end, err);
                if (r != R_FLIP) break;
                // flip: narrow the last 32 KiB of u16 entries into the u8 preamble
                u8_off = ((2 * (uint64_t)WIN + 2 * pos16) + 255) & ~255ull;
                if (u8_off + WIN + 258 > pd.cap) {
                    r = R_OVERFLOW;
                    break;
                }
                const uint16_t *srcw = out16 + pos16 - WIN;
                uint8_t *dw = dst + u8_off;
                for (uint32_t q = lane; q < WIN; q += WAVE) dw[q] = (uint8_t)srcw[q];
                compiler_fence();
                out8 = dw + WIN;
                cap8 = pd.cap - u8_off - WIN;
                pos8 = 0;
                clean_base = pos16;
                mode = M_CLEAN;
            }
        }
        if (r == R_FLIP) {
            // stored block ended a marker run: same flip as above
            u8_off = ((2 * (uint64_t)WIN + 2 * pos16) + 255) & ~255ull;
            if (u8_off + WIN + 258 > pd.cap) {
                r = R_OVERFLOW;
            } else {
                const uint16_t *srcw = out16 + pos16 - WIN;
                uint8_t *dw = dst + u8_off;
                for (uint32_t q = lane; q < WIN; q += WAVE) dw[q] = (uint8_t)srcw[q];
                compiler_fence();
                out8 = dw + WIN;
                cap8 = pd.cap - u8_off - WIN;
                pos8 = 0;
                clean_base = pos16;
                mode = M_CLEAN;
                r = R_EOB;
            }
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
    if (lane == 0) {
        po->end_bit = end_bit;
        po->n16 = pos16;
        po->n8 = pos8;
        po->u8_off = u8_off;
        po->bits_done = br.pos() - pd.start_bit;
        po->status = status;
        po->err = err;
        po->nblocks = nblocks;
        po->pad = (uint32_t)mode;
    }
}

// Persistent inflate kernel: every workgroup (one wavefront) pulls piece
// indices from a device queue until it is drained.  `order` lists pieces
// largest-first so the long pieces start early.
__global__ void __launch_bounds__(64) k_inflate(const uint8_t *__restrict__ in, uint64_t in_len,
                                                const PieceDesc *__restrict__ pieces,
                                                const uint32_t *__restrict__ order, uint32_t n,
                                                uint32_t *queue, PieceOut *out) {
    __shared__ InflateShared sh;
    for (;;) {
        if (lane_id() == 0) sh.qi = atomicAdd(queue, 1u);
        __syncthreads();
        uint32_t qi = uni(sh.qi);
        __syncthreads();
        if (qi >= n) break;
        uint32_t pi = order[qi];
        PieceDesc pd = pieces[pi];
        inflate_piece(sh, in, in_len, pd, out + pi);
        __syncthreads();
    }
}

}  // namespace gz
