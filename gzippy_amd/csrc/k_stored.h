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
__global__ void __launch_bounds__(256) k_stored_scan(const uint8_t *__restrict__ in, uint64_t n, uint64_t first,
                                                     uint64_t last, unsigned long long *__restrict__ cand,
                                                     uint32_t cap, uint32_t *count) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 4;
    for (uint64_t base = first + ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
         base < last; base += stride) {
        // bytes base .. base + 7 cover the headers at base .. base + 3
        const uint64_t w = base >> 2;
        const uint32_t sh = (uint32_t)(base & 3) * 8;
        const uint32_t a = load_word(in, n, w), b = load_word(in, n, w + 1), c2 = load_word(in, n, w + 2);
        const uint64_t lo = ((uint64_t)b << 32 | a) >> sh | (sh ? (uint64_t)c2 << (64 - sh) : 0ull);
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint64_t q = base + k;
            const uint64_t v = lo >> (8 * k);  // bytes q .. q + 4 in the low 40 bits (k <= 3)
            const uint32_t hdr = (uint32_t)v & 0xFF;
            const uint32_t len = (uint32_t)(v >> 8) & 0xFFFF, nlen = (uint32_t)(v >> 24) & 0xFFFF;
            const bool ok = q < last && ((hdr >> 1) & 3) == 0 && (len ^ 0xFFFFu) == nlen && q + 5 + len <= last;
            uint32_t slot;
            const uint64_t m = ballot(ok);
            uint32_t wbase = 0;
            if (m) {
                if (lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1))
                    wbase = atomicAdd(count, (uint32_t)__popcll(m));
                wbase = __shfl(wbase, __ffsll((unsigned long long)m) - 1);
            }
            slot = wbase + mbcnt(m);
            if (ok && slot < cap) cand[slot] = q | (uint64_t)len << 40 | (uint64_t)(hdr & 1) << 56;
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
