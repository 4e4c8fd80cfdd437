// k_blockfinder.h — bit-granular DEFLATE block finder.
//
// Candidate predicate (blockfinder_validation.rs), evaluated independently at
// every bit offset, one lane per offset:
//   dynamic (BTYPE=10): is_deflate_candidate_n on 13 bits (:68-98: BFINAL=0,
//     BTYPE=10, HLIT<=29, HDIST<=29); the bit offset + 17 must lie inside the
//     data (scalar scan break, :693-696); precode leaf count 128, or 64 with
//     no length > 1 (validate_precode :190 + parse_precode/
//     is_valid_huffman_lengths :772-834); then the full code-length decode of
//     validate_huffman_codes (:836-944): repeat rules, is_eof before each
//     read, EOB length > 0, litlen and distance lengths Kraft-complete (or a
//     single 1-bit code).
//   stored (BTYPE=00): find_uncompressed_blocks (:370-470): LEN == ~NLEN at a
//     byte boundary, 3 zero header bits before it, earliest start =
//     byte_bit - trailing zero bits (<= 10), seek = byte_bit - 3, LEN > 0 and
//     the payload fits.
// Stage A tests the cheap header/precode conditions and compacts survivors
// (wave ballot + one atomic per wave); stage B runs the code-length decode
// densely over the survivors, one lane each.
#pragma once
#include "gz_common.h"
#include "gzgpu.h"

namespace gz {

// inverse of the precode transmission order: position of symbol s
__constant__ uint8_t c_pre_pos[19] = {3, 17, 15, 13, 11, 9, 7, 5, 4, 6, 8, 10, 12, 14, 16, 18, 0, 1, 2};

__device__ __forceinline__ void wave_append(bool ok, uint32_t *count, uint32_t &slot) {
    uint64_t m = ballot(ok);
    uint32_t base = 0;
    if (m) {
        if (lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1)) base = atomicAdd(count, (uint32_t)__popcll(m));
        base = __shfl(base, __ffsll((unsigned long long)m) - 1);
    }
    slot = base + mbcnt(m);
}

// Header + precode conditions of the dynamic predicate at `o` (stage A).
__device__ __forceinline__ bool dyn_stage_a(const uint8_t *__restrict__ in, uint64_t n, uint64_t o) {
    uint64_t h = peek64(in, n, o);
    uint32_t hclen = (uint32_t)(h >> 13) & 15;
    bool ok = (h & 7) == 4 /* BFINAL=0, BTYPE=10 */ && ((h >> 3) & 31) <= 29 && ((h >> 8) & 31) <= 29;
    if (!ok) return false;
    uint64_t p = peek64(in, n, o + 17);
    uint32_t leaves = 0, gt1 = 0;
#pragma unroll
    for (uint32_t i = 0; i < 19; i++) {
        uint32_t l = (uint32_t)(p >> (3 * i)) & 7;
        if (i >= hclen + 4) l = 0;
        leaves += l ? (1u << (7 - l)) : 0u;
        gt1 += l > 1;
    }
    return leaves == 128 || (leaves == 64 && gt1 == 0);
}

// Leaf-sum table for the precode test: entry v (four 3-bit code lengths,
// first in the low bits) = sum of 2^(7-l) over its nonzero lengths, plus the
// count of lengths > 1 in bits 16 and up.  Five lookups cover the 19 lengths.
__device__ __forceinline__ uint32_t precode_lut_entry(uint32_t v) {
    uint32_t leaves = 0, gt1 = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t l = (v >> (3 * i)) & 7;
        leaves += l ? (1u << (7 - l)) : 0u;
        gt1 += l > 1;
    }
    return leaves | gt1 << 16;
}

// Precode condition of the dynamic predicate, from the 57 bits after the 17
// header bits (`p`) and HCLEN: leaf count 128, or 64 with no length > 1
// (lengths past HCLEN + 4 count as 0).
__device__ __forceinline__ bool precode_ok_lut(const uint32_t *lut, uint64_t p, uint32_t hclen) {
    const uint32_t nb = 3 * (hclen + 4);  // 12 .. 57
    p &= (1ull << nb) - 1;
    const uint32_t s = lut[(uint32_t)p & 0xFFF] + lut[(uint32_t)(p >> 12) & 0xFFF] + lut[(uint32_t)(p >> 24) & 0xFFF] +
                       lut[(uint32_t)(p >> 36) & 0xFFF] + lut[(uint32_t)(p >> 48) & 0x1FF];
    const uint32_t leaves = s & 0xFFFF, gt1 = s >> 16;
    return leaves == 128 || (leaves == 64 && gt1 == 0);
}

// The stored-block conditions at byte-aligned offset o whose LEN/NLEN word
// matched; fills the hit fields (find_uncompressed_blocks,
// blockfinder_validation.rs:370-470).
__device__ __forceinline__ bool stored_hit(const uint8_t *__restrict__ in, uint64_t n, uint64_t o, uint64_t start_bit,
                                           uint64_t end_bit, uint64_t sb, uint64_t until, uint32_t size,
                                           uint64_t &earliest, uint64_t &latest, bool &succ) {
    const uint64_t by = o >> 3;
    if (!(by >= sb && by < until && o >= 10)) return false;
    const uint64_t first = o - 10;
    const uint32_t raw = (uint32_t)in[first >> 3] | (uint32_t)in[(first >> 3) + 1] << 8;
    const uint32_t pre = (raw >> (first & 7)) & 0x3FF;
    if ((pre & (7u << 7)) != 0) return false;
    uint32_t tz = 3;
    for (uint32_t j = 4; j <= 10; j++) {
        if (pre & (1u << (10 - j))) break;
        tz = j;
    }
    earliest = o - tz;
    latest = o - 3;
    const uint32_t len = size & 0xFFFF;
    if (!(earliest >= start_bit && latest < end_bit && len > 0 && by + 4 + len <= n)) return false;
    // successor header after the payload (planning hint only)
    succ = false;
    const uint64_t nb = by + 4 + len;
    if (nb < n) {
        const uint32_t h3 = in[nb] & 7, bt = h3 >> 1;
        if (bt == 0) {
            if (nb + 5 <= n) {
                const uint32_t s2 = (uint32_t)in[nb + 1] | (uint32_t)in[nb + 2] << 8 | (uint32_t)in[nb + 3] << 16 |
                                    (uint32_t)in[nb + 4] << 24;
                succ = ((s2 ^ (s2 >> 16)) & 0xFFFF) == 0xFFFF;
            }
        } else if (bt == 2) {
            // a final dynamic block would fail BFINAL=0: test the header
            // with BFINAL masked off
            succ = (in[nb] & 1) ? ((peek64(in, n, nb * 8) >> 3 & 31) <= 29) : dyn_stage_a(in, n, nb * 8);
        }
    }
    return true;
}

// One scan range of stage A: candidate offsets in [start_bit, end_bit);
// chunk0 is the global index of its first 32-offset chunk (ranges are laid
// end to end in chunk space, so one launch covers every member of a batch).
// final_ok: dynamic headers with BFINAL = 1 are reported too (the planner's
// scan of each member's last stretch, where the final block starts; the
// reference predicate, and every other scan, takes BFINAL = 0 only).
struct BfRange {
    uint64_t start_bit, end_bit, chunk0;
    uint32_t final_ok, pad;
};

// Stage A.  Each lane owns 32 consecutive bit offsets [o0, o0 + 32) of one
// range and reads their 128 bits once into registers.  The 13-bit header test
// runs on all 32 offsets at once as bit operations on a 64-bit word; the
// precode leaf sum (five lookups in an LDS table) only on the offsets that
// pass it (about one in nine); the stored-block LEN/NLEN test on the 4
// byte-aligned offsets, the rest of the stored conditions only on a match.
// Survivors collect in a per-wave LDS buffer and go to HBM with one atomic
// per 512 (one counter shared by every wave is the contended resource; an
// atomic per wave and iteration made the kernel several times slower).
// Hits are appended with one atomic per wave.
__global__ void __launch_bounds__(256) k_bf_stage_a(const uint8_t *__restrict__ in, uint64_t n,
                                                    const BfRange *__restrict__ ranges, uint32_t nranges,
                                                    uint64_t nchunk, uint64_t *__restrict__ surv, uint32_t surv_cap,
                                                    uint32_t *surv_count, gzgpu_boundary *__restrict__ hits,
                                                    uint32_t hit_cap, uint32_t *hit_count) {
    constexpr uint32_t WBUF = 512;
    __shared__ uint32_t pre_lut[4096];
    __shared__ uint64_t wbuf[4][WBUF];  // blockDim.x == 256: four wavefronts
    for (uint32_t v = threadIdx.x; v < 4096; v += blockDim.x) pre_lut[v] = precode_lut_entry(v);
    __syncthreads();
    uint64_t *mybuf = wbuf[threadIdx.x / WAVE];
    uint32_t wcnt = 0;  // wave-uniform fill of mybuf
    auto flush = [&]() {
        compiler_fence();
        uint32_t base_slot = 0;
        if (lane_id() == 0) base_slot = atomicAdd(surv_count, wcnt);
        base_slot = (uint32_t)__builtin_amdgcn_readfirstlane((int)base_slot);
        for (uint32_t i = lane_id(); i < wcnt; i += WAVE)
            if (base_slot + i < surv_cap) surv[base_slot + i] = mybuf[i];
        compiler_fence();
        wcnt = 0;
    };
    const uint64_t total = n * 8;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // wave-aligned iterations so every lane reaches the ballots
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nchunk; base += stride) {
        const uint64_t c = base + threadIdx.x;
        const bool live = c < nchunk;
        // this lane's range: the last one whose chunk0 <= c
        uint32_t lo_r = 0, hi_r = nranges;
        while (hi_r - lo_r > 1) {
            const uint32_t mid = (lo_r + hi_r) >> 1;
            if (ranges[mid].chunk0 <= c) lo_r = mid;
            else hi_r = mid;
        }
        const uint64_t start_bit = ranges[lo_r].start_bit, end_bit = ranges[lo_r].end_bit;
        const bool final_ok = ranges[lo_r].final_ok != 0;
        const uint64_t o0 = (start_bit & ~31ull) + 32 * (c - ranges[lo_r].chunk0);
        const uint64_t sb = ((start_bit + 3 + 7) / 8) > 1 ? (start_bit + 3 + 7) / 8 : 1;
        uint64_t until = (end_bit + 7) / 8;
        if (n < 4) until = 0;
        else if (until > n - 4) until = n - 4;
        // dynamic candidates need o + 17 < total
        const uint64_t dyn_end = total > 17 ? (end_bit < total - 17 ? end_bit : total - 17) : 0;
        uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
        if (live) {
            const uint64_t wi = o0 >> 5;
            if ((wi + 4) * 4 <= n) {
                gptr<const uint32_t> w = (gptr<const uint32_t>)to_global(in) + wi;
                w0 = w[0];
                w1 = w[1];
                w2 = w[2];
                w3 = w[3];
            } else {
                w0 = load_word(in, n, wi);
                w1 = load_word(in, n, wi + 1);
                w2 = load_word(in, n, wi + 2);
                w3 = load_word(in, n, wi + 3);
            }
        }
        // ---- header test on all 32 offsets: bit j of the mask is the test
        // at o0 + j.  BFINAL = 0 (either, in a final_ok range) and BTYPE = 10
        // (bits 0, 1, 2 = 0, 0, 1);
        // HLIT <= 29 and HDIST <= 29 (not all of bits 4-7, not all of 9-12)
        const uint64_t X = (uint64_t)w0 | (uint64_t)w1 << 32;
        uint32_t hm = (uint32_t)((final_ok ? ~0ull : ~X) & ~(X >> 1) & (X >> 2) & ~((X >> 4) & (X >> 5) & (X >> 6) & (X >> 7)) &
                                 ~((X >> 9) & (X >> 10) & (X >> 11) & (X >> 12)));
        // offsets inside [start_bit, dyn_end)
        {
            const uint64_t lo = start_bit > o0 ? start_bit - o0 : 0;
            const uint64_t hi = dyn_end > o0 ? dyn_end - o0 : 0;
            const uint32_t lo32 = lo > 32 ? 32u : (uint32_t)lo, hi32 = hi > 32 ? 32u : (uint32_t)hi;
            const uint32_t below_hi = hi32 >= 32 ? 0xFFFFFFFFu : ((1u << hi32) - 1);
            const uint32_t below_lo = lo32 >= 32 ? 0xFFFFFFFFu : ((1u << lo32) - 1);
            hm &= below_hi & ~below_lo;
            if (!live) hm = 0;
        }
        // ---- precode test on the header survivors -> survivor mask
        uint32_t sm = 0;
        for (uint32_t hh = hm; hh; hh &= hh - 1) {
            const uint32_t j = (uint32_t)__builtin_ctz(hh);
            const uint32_t h = __builtin_amdgcn_alignbit(w1, w0, j);
            const uint32_t q = j + 17, k = q >> 5, r = q & 31;
            const uint32_t a = k ? w1 : w0, b = k ? w2 : w1, d = k ? w3 : w2;
            const uint64_t p =
                (uint64_t)__builtin_amdgcn_alignbit(b, a, r) | (uint64_t)__builtin_amdgcn_alignbit(d, b, r) << 32;
            if (precode_ok_lut(pre_lut, p, (h >> 13) & 15)) sm |= 1u << j;
        }
        // survivors: lanes write after a prefix sum, into the wave's LDS
        // buffer (flushed when it would overflow), or straight to HBM with
        // their own atomic when one iteration alone has more than it holds
        {
            const uint32_t cnt = (uint32_t)__builtin_popcount(sm);
            const uint32_t incl = wave_scan_incl(cnt);
            const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
            if (tot > WBUF) {
                uint32_t base_slot = 0;
                if (lane_id() == 0) base_slot = atomicAdd(surv_count, tot);
                base_slot = (uint32_t)__builtin_amdgcn_readfirstlane((int)base_slot);
                uint32_t slot = base_slot + incl - cnt;
                for (uint32_t ss = sm; ss; ss &= ss - 1, slot++)
                    if (slot < surv_cap) surv[slot] = o0 + (uint32_t)__builtin_ctz(ss);
            } else if (tot) {
                if (wcnt + tot > WBUF) flush();
                uint32_t slot = wcnt + incl - cnt;
                for (uint32_t ss = sm; ss; ss &= ss - 1, slot++) mybuf[slot] = o0 + (uint32_t)__builtin_ctz(ss);
                wcnt += tot;
            }
        }
        // ---- stored blocks at the 4 byte-aligned offsets
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint64_t o = o0 + 8 * k;
            const uint32_t size = __builtin_amdgcn_alignbit(w1, w0, 8 * k);
            bool cand = live && o < end_bit && ((size ^ (size >> 16)) & 0xFFFF) == 0xFFFF;
            if (!ballot(cand)) continue;
            uint64_t earliest = 0, latest = 0;
            bool succ = false, st = false;
            if (cand) st = stored_hit(in, n, o, start_bit, end_bit, sb, until, size, earliest, latest, succ);
            uint32_t slot;
            wave_append(st, hit_count, slot);
            if (st && slot < hit_cap) {
                hits[slot].bit_offset = earliest;
                hits[slot].seek_bit = latest;
                hits[slot].kind = 0;
                hits[slot].info = (succ ? 1u : 0u) | (size & 0xFFFFu) << 16;
            }
        }
    }
    if (wcnt) flush();
}

// Full dynamic-header validation of one survivor (one lane).
// info of a dynamic hit: HLIT bits 0-4, HDIST 8-12, HCLEN 16-19, BFINAL bit 7
// (set only by final_ok scans), the planner's estimate in bits 20-31
constexpr uint32_t HIT_FINAL_SHIFT = 7;
__device__ bool validate_dynamic(const uint8_t *__restrict__ in, uint64_t n, uint64_t o, uint8_t *lut,
                                 uint32_t &info) {
    const uint64_t total = n * 8;
    uint64_t h = peek64(in, n, o);
    uint32_t hlit = (uint32_t)(h >> 3) & 31, hdist = (uint32_t)(h >> 8) & 31, hclen = (uint32_t)(h >> 13) & 15;
    info = hlit | hdist << 8 | hclen << 16 | ((uint32_t)h & 1u) << HIT_FINAL_SHIFT;
    uint64_t p = peek64(in, n, o + 17);
    uint32_t pc = hclen + 4;
    // canonical precode: counts, then codes assigned by symbol order
    uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (uint32_t i = 0; i < 19; i++) {
        uint32_t l = i < pc ? (uint32_t)(p >> (3 * i)) & 7 : 0;
#pragma unroll
        for (uint32_t L = 1; L < 8; L++) cnt[L] += l == L;
    }
    uint32_t next[8];
    uint32_t code = 0;
    next[0] = 0;
#pragma unroll
    for (uint32_t L = 1; L < 8; L++) {
        code = (code + (L > 1 ? cnt[L - 1] : 0)) << 1;
        next[L] = code;
    }
    uint32_t *lw = reinterpret_cast<uint32_t *>(lut);
#pragma unroll
    for (int i = 0; i < 32; i++) lw[i] = 0;
#pragma unroll
    for (uint32_t s = 0; s < 19; s++) {
        uint32_t ps = c_pre_pos[s];
        uint32_t l = ps < pc ? (uint32_t)(p >> (3 * ps)) & 7 : 0;
        if (l) {
            uint32_t c = 0;
#pragma unroll
            for (uint32_t L = 1; L < 8; L++)
                if (l == L) c = next[L]++;
            uint32_t rev = __brev(c) >> (32 - l);
            for (uint32_t f = rev; f < 128; f += 1u << l) lut[f] = (uint8_t)(s | l << 5);
        }
    }
    uint64_t pos = o + 17 + 3 * pc;
    const uint32_t hl = hlit + 257, total_codes = hl + hdist + 1;
    uint32_t i = 0, prev = 0, len256 = 0;
    uint32_t sum_ll = 0, sum_d = 0, gt1_ll = 0, gt1_d = 0;
    // Streaming bit buffer: `buf` holds the nb >= 32 bits from pos on and the
    // next input word is already loaded, so the dependent load of each step
    // (a code-length symbol takes at most 14 bits) is a word ahead instead of
    // a full read per symbol.
    uint64_t wi = pos >> 5;
    uint64_t buf = ((uint64_t)load_word(in, n, wi) | (uint64_t)load_word(in, n, wi + 1) << 32) >> (pos & 31);
    uint32_t nb = 64 - (uint32_t)(pos & 31);
    uint32_t nxt = load_word(in, n, wi + 2);
    wi += 3;
    while (i < total_codes) {
        if (pos >= total) return false;  // is_eof
        if (nb < 32) {
            buf |= (uint64_t)nxt << nb;
            nb += 32;
            nxt = load_word(in, n, wi++);
        }
        uint64_t w = buf;
        uint32_t e = lut[w & 127];
        uint32_t L = e >> 5, sym = e & 31;
        if (L == 0) return false;
        uint32_t used = L;
        w >>= L;
        uint32_t val, rep;
        if (sym < 16) {
            val = sym;
            rep = 1;
        } else if (sym == 16) {
            if (i == 0) return false;
            val = prev;
            rep = 3 + (uint32_t)(w & 3);
            used += 2;
        } else if (sym == 17) {
            val = 0;
            rep = 3 + (uint32_t)(w & 7);
            used += 3;
        } else {
            val = 0;
            rep = 11 + (uint32_t)(w & 127);
            used += 7;
        }
        pos += used;
        buf >>= used;
        nb -= used;
        if (i + rep > total_codes) return false;
        if (i <= 256 && 256 < i + rep) len256 = val;
        uint32_t n_ll = i < hl ? (hl - i < rep ? hl - i : rep) : 0;
        uint32_t n_d = rep - n_ll;
        if (val) {
            uint32_t c = 1u << (15 - val);
            sum_ll += n_ll * c;
            sum_d += n_d * c;
            if (val > 1) {
                gt1_ll += n_ll;
                gt1_d += n_d;
            }
        }
        prev = val;
        i += rep;
        if (sum_ll > 32768 || sum_d > 32768) return false;  // over-subscribed: can only grow
    }
    if (len256 == 0) return false;
    bool ok_ll = sum_ll == 32768 || (sum_ll == 16384 && gt1_ll == 0);
    bool ok_d = sum_d == 32768 || (sum_d == 16384 && gt1_d == 0);
    return ok_ll && ok_d;
}

// Survivor count from stage A's counter (capped at the buffer), so that the
// host need not read it between the stages.
__global__ void __launch_bounds__(256) k_bf_stage_b(const uint8_t *__restrict__ in, uint64_t n,
                                                    const uint64_t *__restrict__ surv, const uint32_t *surv_count,
                                                    uint32_t surv_cap, gzgpu_boundary *__restrict__ hits,
                                                    uint32_t hit_cap, uint32_t *hit_count) {
    __shared__ __align__(16) uint8_t luts[256][128];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t counted = (uint32_t)__builtin_amdgcn_readfirstlane((int)*surv_count);
    const uint32_t nsurv = counted < surv_cap ? counted : surv_cap;
    for (uint32_t base = blockIdx.x * blockDim.x; base < nsurv; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool ok = false;
        uint64_t o = 0;
        uint32_t info = 0;
        if (i < nsurv) {
            o = surv[i];
            ok = validate_dynamic(in, n, o, luts[threadIdx.x], info);
        }
        uint32_t slot;
        wave_append(ok, hit_count, slot);
        if (ok && slot < hit_cap) {
            hits[slot].bit_offset = o;
            hits[slot].seek_bit = o;
            hits[slot].kind = 2;
            hits[slot].info = info;
        }
    }
}

// Output estimate of every dynamic-header hit, in bits 20-31 of its info:
// the block's expected output per compressed bit, as its code lengths imply
// it (symbol s drawn with probability 2^-len(s)): entries and bits of an
// average literal/length symbol, a length symbol's distance bits averaged
// the same way.  A planning estimate (cutting a member into pieces of equal
// decode time), not a parity value: log2(entries per bit) in 1/256 steps,
// offset by 8, clamped to 12 bits.
//
// One wavefront per hit, so the serial walk over the code lengths (a
// hundred or more precode symbols) becomes ~15 rounds: every lane decodes
// the precode symbol that would start at bit P + lane from a 128-bit window,
// a readlane walk from lane 0 picks the true starts, a prefix sum gives each
// its code index, and each taken lane adds its run's weight.
__global__ void __launch_bounds__(64) k_bf_estimate(const uint8_t *__restrict__ in, uint64_t n,
                                                    gzgpu_boundary *__restrict__ hits, const uint32_t *hit_count,
                                                    uint32_t hit_cap) {
    __shared__ uint8_t lut[128];
    const uint32_t lane = lane_id();
    const uint32_t counted = uni(*hit_count);
    const uint32_t nh = counted < hit_cap ? counted : hit_cap;
    for (uint32_t hidx = blockIdx.x; hidx < nh; hidx += gridDim.x) {
        if (uni(hits[hidx].kind) != 2) continue;
        const uint64_t o = uni64(hits[hidx].bit_offset);
        const uint64_t h = uni64(peek64(in, n, o));
        const uint32_t hl = ((uint32_t)(h >> 3) & 31) + 257, hdist = (uint32_t)(h >> 8) & 31,
                       pc = ((uint32_t)(h >> 13) & 15) + 4;
        const uint64_t pre = uni64(peek64(in, n, o + 17));
        // precode LUT: symbol s = lane, canonical codes in symbol order
        uint32_t l = 0;
        if (lane < 19) {
            const uint32_t ps = c_pre_pos[lane];
            l = ps < pc ? (uint32_t)(pre >> (3 * ps)) & 7 : 0;
        }
        uint32_t code = 0, cprev = 0, my_code = 0;
#pragma unroll
        for (uint32_t L = 1; L < 8; L++) {
            const uint64_t mL = ballot(l == L);
            code = (code + cprev) << 1;
            if (l == L) my_code = code + mbcnt(mL);
            cprev = (uint32_t)__popcll(mL);
        }
        for (uint32_t f = lane; f < 128; f += 64) lut[f] = 0;
        __syncthreads();
        if (l) {
            const uint32_t rev = __brev(my_code) >> (32 - l);
            for (uint32_t f = rev; f < 128; f += 1u << l) lut[f] = (uint8_t)(lane | l << 5);
        }
        __syncthreads();
        const uint32_t total = hl + hdist + 1;
        uint64_t P = o + 17 + 3 * pc;
        uint32_t i = 0, prev = 0;
        float e_ent = 0.f, e_bits = 0.f, m_len = 0.f, d_bits = 0.f, d_w = 0.f;
        for (uint32_t round = 0; i < total && round < 256; round++) {
            const uint64_t W0 = uni64(peek64(in, n, P)), W1 = uni64(peek64(in, n, P + 64));
            const uint64_t ch = lane ? (W0 >> lane) | (W1 << (64 - lane)) : W0;
            const uint32_t e = lut[ch & 127];
            const uint32_t L = e >> 5, sym = e & 31;
            const uint64_t x = ch >> L;
            const uint32_t ext = sym < 16 ? 0u : sym == 16 ? 2u : sym == 17 ? 3u : 7u;
            const uint32_t rep = sym < 16    ? 1u
                                 : sym == 16 ? 3u + (uint32_t)(x & 3)
                                 : sym == 17 ? 3u + (uint32_t)(x & 7)
                                             : 11u + (uint32_t)(x & 127);
            const uint32_t nxt = lane + L + ext;
            const uint64_t OK = ballot(L != 0);
            if (!(OK & 1)) break;  // (not for a validated header)
            uint64_t M = 0;
            uint32_t st = 0;
            for (;;) {
                M |= 1ull << st;
                const uint32_t t = uni((uint32_t)__builtin_amdgcn_readlane((int)nxt, (int)st));
                st = t;
                if (t >= 64 || !((OK >> t) & 1)) break;
            }
            M = uni64(M);
            const bool inM = (M >> lane) & 1;
            const uint32_t v = inM ? rep : 0u;
            const uint32_t incl = wave_scan_incl(v);
            const uint32_t idx = i + incl - v;
            const bool take = inM && idx < total;
            const uint64_t T = uni64(ballot(take));
            // a repeat (16) takes the value of the nearest earlier taken
            // symbol that is not a repeat, or the carried one
            const uint64_t NZ = T & ballot(sym != 16);
            const uint64_t below = NZ & ((1ull << lane) - 1);
            const int src = below ? 63 - __builtin_clzll(below) : 0;
            const uint32_t def = sym < 16 ? sym : 0u;
            const uint32_t from = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)def);
            const uint32_t val = sym == 16 ? (below ? from : prev) : def;
            if (take && val) {
                const float w = (float)(1u << (15 - val));
                const uint32_t hi = idx + rep < total ? idx + rep : total;
                const uint32_t lo = idx;
                const uint32_t ll_hi = hi < hl ? hi : hl;
                if (ll_hi > lo) {
                    const uint32_t lit_hi = ll_hi < 256 ? ll_hi : 256;
                    if (lit_hi > lo) {
                        e_ent += w * (float)(lit_hi - lo);
                        e_bits += w * (float)(val * (lit_hi - lo));
                    }
                    if (lo <= 256 && 256 < ll_hi) e_bits += w * (float)val;  // end-of-block
                    for (uint32_t q = lo > 257 ? lo : 257; q < ll_hi && q < 286; q++) {
                        const uint32_t k = q - 257, E = c_lextra[k];
                        e_ent += w * ((float)c_lbase[k] + 0.5f * (float)((1u << E) - 1));
                        e_bits += w * (float)(val + E);
                        m_len += w;
                    }
                }
                for (uint32_t d = (lo > hl ? lo : hl); d < hi && d - hl < 30; d++) {
                    d_w += w;
                    d_bits += w * (float)(val + c_dextra[d - hl]);
                }
            }
            if (!T) break;
            const uint32_t last = uni((uint32_t)(63 - __builtin_clzll(T)));
            prev = uni((uint32_t)__builtin_amdgcn_readlane((int)val, (int)last));
            i += uni((uint32_t)__builtin_amdgcn_readlane((int)(incl - v + (take ? rep : 0u)), (int)last));
            i = uni(i);
            P += st;
            P = uni64(P);
            if (T != M) break;  // the header ended inside this round
        }
        for (uint32_t off = 32; off; off >>= 1) {
            e_ent += __shfl_xor(e_ent, off);
            e_bits += __shfl_xor(e_bits, off);
            m_len += __shfl_xor(m_len, off);
            d_bits += __shfl_xor(d_bits, off);
            d_w += __shfl_xor(d_w, off);
        }
        if (lane == 0) {
            const float bits = e_bits + m_len * (d_w > 0.f ? d_bits / d_w : 5.f);
            const float ratio = bits > 0.f && e_ent > 0.f ? e_ent / bits : 1.f;
            int r = (int)((__log2f(ratio) + 8.f) * 256.f + 0.5f);
            r = r < 0 ? 0 : r > 4095 ? 4095 : r;
            hits[hidx].info = (hits[hidx].info & 0xFFFFFu) | (uint32_t)r << 20;
        }
        __syncthreads();
    }
}

}  // namespace gz
