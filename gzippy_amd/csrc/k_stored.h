// k_stored.h — non-speculative split of members made of stored blocks
// (StoredParallel: decompress_stored_parallel, stored_split.rs:265).
//
// In a run of stored blocks every block header sits at bit 0 of a byte (a
// stored block ends on a byte boundary), its LEN/NLEN in the four bytes
// after it and its payload next (walk_stored_chain, stored_split.rs:195).
// So the chain needs no bit-level decode:
//   k_stored_scan   every byte q of the member is tested as such a header:
//                   BTYPE bits (1-2) zero, LEN == ~NLEN, payload inside the
//                   member; candidates (a few per MiB of random data besides
//                   the true ones) are appended to a list;
//   host            walks the chain from the member's first byte through the
//                   candidates: q -> q + 5 + LEN, until a BFINAL block that
//                   ends exactly where the trailer begins (else the member is
//                   not pure-stored and takes the speculative path);
//   k_stored_copy   copies every block's payload to its output offset, one
//                   workgroup per block, 4-byte stores at 4-byte-aligned
//                   output addresses fed by realigned dword loads.
// The CRC32 of the output is then folded like any other member's.
#pragma once
#include "gz_common.h"

namespace gz {

// candidate: byte offset (bits 0-39), LEN (40-55), BFINAL (56)
//
// Each thread tests SCAN_B consecutive header positions per step from six
// aligned words.  LEN == ~NLEN (the 32 bits after the header byte, one
// v_alignbyte, a shift and an xor per position) is the rare condition and is
// tested first; the header bits, the bounds and the append run only for the
// positions that pass it (a few per MiB besides the true headers), so the
// scan is bound by its loads rather than by a ballot per byte.
constexpr uint32_t SCAN_B = 16;
__global__ void __launch_bounds__(256) k_stored_scan(const uint8_t *__restrict__ in, uint64_t n, uint64_t first,
                                                     uint64_t last, unsigned long long *__restrict__ cand,
                                                     uint32_t cap, uint32_t *count) {
    const uint64_t a0 = first & ~(uint64_t)3;  // word-aligned start
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * SCAN_B;
    for (uint64_t base = a0 + ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * SCAN_B; base < last;
         base += stride) {
        // bytes base .. base + 23 cover the headers at base .. base + 15 and
        // their LEN/NLEN
        const uint64_t w = base >> 2;
        uint32_t x[SCAN_B / 4 + 2];
#pragma unroll
        for (uint32_t i = 0; i < SCAN_B / 4 + 2; i++) x[i] = load_word(in, n, w + i);
        uint32_t m = 0;
#pragma unroll
        for (uint32_t k = 0; k < SCAN_B; k++) {
            const uint32_t j = k + 1;  // byte of LEN's low half, from base
            const uint32_t t = (j & 3) ? __builtin_amdgcn_alignbyte(x[j / 4 + 1], x[j / 4], j & 3) : x[j / 4];
            m |= ((t ^ (t >> 16)) & 0xFFFFu) == 0xFFFFu ? 1u << k : 0u;
        }
        if (ballot(m != 0)) {
            for (uint32_t k = 0; k < SCAN_B; k++) {
                if (!((m >> k) & 1)) continue;
                const uint64_t q = base + k;
                const uint32_t hdr = (x[k / 4] >> (8 * (k & 3))) & 0xFF;
                const uint32_t j = k + 1;
                const uint32_t t = (j & 3) ? __builtin_amdgcn_alignbyte(x[j / 4 + 1], x[j / 4], j & 3) : x[j / 4];
                const uint32_t len = t & 0xFFFF;
                if (q < first || q >= last || ((hdr >> 1) & 3) != 0 || q + 5 + len > last) continue;
                const uint32_t slot = atomicAdd(count, 1u);
                if (slot < cap) cand[slot] = q | (uint64_t)len << 40 | (uint64_t)(hdr & 1) << 56;
            }
        }
    }
}

struct StoredRun {
    uint64_t src;  // input byte offset of the payload
    uint64_t dst;  // output address
    uint32_t len;
    uint32_t pad;
};

// One workgroup per run (grid-stride over runs).  The input is read with
// aligned dword loads realigned by v_alignbyte; loads that would reach past
// the input fall back to bounds-checked ones.
__global__ void __launch_bounds__(256) k_stored_copy(const uint8_t *__restrict__ in, uint64_t n,
                                                     const StoredRun *__restrict__ runs, uint32_t nruns) {
    gptr<const uint32_t> inw = (gptr<const uint32_t>)to_global(in);
    for (uint32_t r = blockIdx.x; r < nruns; r += gridDim.x) {
        const StoredRun R = runs[r];
        gptr<uint8_t> o8 = to_global((uint8_t *)(uintptr_t)R.dst);
        const uint64_t d0 = 0, s0 = R.src;
        // head: bytes until the output is 4-byte aligned
        const uint32_t mis = (uint32_t)((4 - (R.dst & 3)) & 3);
        const uint32_t head = mis < R.len ? mis : R.len;
        if (threadIdx.x < head) o8[d0 + threadIdx.x] = in[s0 + threadIdx.x];
        const uint32_t body = (R.len - head) & ~3u;
        gptr<uint32_t> dw = (gptr<uint32_t>)(o8 + d0 + head);
        const uint64_t s = s0 + head;
        const uint32_t sh = (uint32_t)(s & 3);
        const uint64_t w0 = s >> 2;
        const uint64_t nw = n >> 2;  // whole words inside the input
        for (uint32_t i = threadIdx.x; i < body / 4; i += blockDim.x) {
            const uint64_t wi = w0 + i;
            uint32_t v;
            if (wi + 1 < nw) {
                v = sh ? __builtin_amdgcn_alignbyte(inw[wi + 1], inw[wi], sh) : inw[wi];
            } else {
                const uint32_t lo = load_word(in, n, wi), hi = load_word(in, n, wi + 1);
                v = sh ? __builtin_amdgcn_alignbyte(hi, lo, sh) : lo;
            }
            dw[i] = v;
        }
        const uint32_t tail = R.len - head - body;
        if (threadIdx.x < tail) o8[d0 + head + body + threadIdx.x] = in[s0 + head + body + threadIdx.x];
    }
}

}  // namespace gz
