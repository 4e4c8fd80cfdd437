// k_crc.h — gzip CRC32 (reflected 0xEDB88320) on the device.
//
// Stage 1 (k_crc_segments): every workgroup folds one segment of <= 64 KiB;
// each of its 256 threads folds the segment's 16-byte blocks 4096 bytes
// apart (two slicing-by-8 steps per block, from a zero register), and the
// partial CRCs combine with the GF(2) shift
// crc(A||B) = crc(A) * x^(8|B|) mod P ^ crc(B) (combine_crc32,
// crc32.rs:127-133; polynomial_multiply_modulo :64; x_power_modulo :104).
// Each segment's CRC leaves stage 1 already shifted over the member bytes
// after it (seg_tail), so stage 2 (k_crc_members) only XORs a member's
// segment values and applies the init/final inversion:
// crc = ~(x^(8D) * 0xFFFFFFFF ^ raw).
#pragma once
#include "gz_common.h"

namespace gz {

constexpr uint32_t CRC_POLY = 0xEDB88320u;
constexpr uint32_t CRC_SEG = 65536;

__device__ __forceinline__ uint32_t gf2_mulmod(uint32_t a, uint32_t b) {
    uint32_t r = 0;
#pragma unroll
    for (int i = 31; i >= 0; i--) {
        r ^= b & (0u - ((a >> i) & 1u));
        b = (b >> 1) ^ (CRC_POLY & (0u - (b & 1u)));
    }
    return r;
}

// c_xbyte[k][j] = x^(8 * j * 16^k) mod P (reflected), filled by the host:
// shifting over n zero bytes costs one multiply per non-zero hex digit of n.
__constant__ uint32_t c_xbyte[16][16];

// x^(8n) mod P
__device__ __forceinline__ uint32_t xpow_bytes(uint64_t n) {
    uint32_t p = 1u << 31;
    for (int k = 0; n; k++, n >>= 4)
        if (n & 15) p = gf2_mulmod(c_xbyte[k][n & 15], p);
    return p;
}


// shift a raw CRC register over `len` zero bytes
__device__ __forceinline__ uint32_t crc_shift(uint32_t c, uint64_t len) { return len ? gf2_mulmod(xpow_bytes(len), c) : c; }

// undo the shift over n zero bytes: divide by x^(8n) mod P.  One step of
// multiplication by x is c -> (c >> 1) ^ (P if c & 1); P has bit 31 set, so
// the top bit of the product tells the low bit of c, and the step inverts.
__device__ __forceinline__ uint32_t crc_unshift(uint32_t c, uint32_t n) {
    for (uint32_t i = 0; i < 8 * n; i++) c = (c & 0x80000000u) ? (((c ^ CRC_POLY) << 1) | 1u) : (c << 1);
    return c;
}


// Raw CRC (from a zero register) of 16 bytes given as four little-endian
// dwords: two slicing-by-8 steps over the tables t[8][256].
__device__ __forceinline__ uint32_t crc16b(const uint32_t (*t)[256], uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) {
    uint32_t c = t[7][w0 & 0xFF] ^ t[6][(w0 >> 8) & 0xFF] ^ t[5][(w0 >> 16) & 0xFF] ^ t[4][w0 >> 24] ^ t[3][w1 & 0xFF] ^
                 t[2][(w1 >> 8) & 0xFF] ^ t[1][(w1 >> 16) & 0xFF] ^ t[0][w1 >> 24];
    const uint32_t lo = c ^ w2;
    return t[7][lo & 0xFF] ^ t[6][(lo >> 8) & 0xFF] ^ t[5][(lo >> 16) & 0xFF] ^ t[4][lo >> 24] ^ t[3][w3 & 0xFF] ^
           t[2][(w3 >> 8) & 0xFF] ^ t[1][(w3 >> 16) & 0xFF] ^ t[0][w3 >> 24];
}

// c * x^(8 * 4096) mod P through the byte tables mx[4][256] (mx[k][b] is the
// product for byte b in position k of c; the map is linear over GF(2)).
__device__ __forceinline__ uint32_t mul_x4k(const uint32_t (*mx)[256], uint32_t c) {
    return mx[0][c & 0xFF] ^ mx[1][(c >> 8) & 0xFF] ^ mx[2][(c >> 16) & 0xFF] ^ mx[3][c >> 24];
}

// Segments: seg_off[i] bytes into `base`, seg_len[i] <= CRC_SEG bytes each,
// one workgroup per segment, folded the way k_resolve_body folds its tiles:
// the 16-byte-aligned blocks [A0, A1) of the address space that hold segment
// bytes are read with coalesced 16-byte loads (bytes outside the segment
// masked to zero), thread t takes blocks t, t + 256, ... (4096 bytes apart)
// and folds them in order with acc = acc * x^(8 * 4096) ^ crc(block), then
// shifts once to A1; the threads' values XOR together.  Zero bytes before the
// segment leave a raw CRC register at 0; the zeros after it (A1 - end) are
// divided out, and the value is shifted over the member bytes after the
// segment (seg_tail).  Only the CRC tables sit in LDS (12 KiB), so many
// workgroups share a CU and keep their loads in flight.
__global__ void __launch_bounds__(256) k_crc_segments(const uint8_t *__restrict__ base, const uint64_t *__restrict__ seg_off,
                                                      const uint32_t *__restrict__ seg_len,
                                                      const uint64_t *__restrict__ seg_tail, uint32_t nseg,
                                                      const uint32_t *__restrict__ tables, uint32_t *__restrict__ seg_crc) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    __shared__ uint32_t t8[8][256];
    __shared__ uint32_t mx[4][256];
    __shared__ uint32_t red;
    for (uint32_t i = threadIdx.x; i < 8 * 256; i += blockDim.x) t8[i >> 8][i & 255] = tables[i];
    for (uint32_t i = threadIdx.x; i < 4 * 256; i += blockDim.x) mx[i >> 8][i & 255] = tables[8 * 256 + i];
    __syncthreads();
    for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        if (threadIdx.x == 0) red = 0;
        __syncthreads();
        const uintptr_t p = (uintptr_t)(base + seg_off[s]), e = p + seg_len[s];
        const uintptr_t A0 = p & ~(uintptr_t)15, A1 = (e + 15) & ~(uintptr_t)15;
        uint32_t acc = 0;
        uintptr_t my_end = 0;  // end of this thread's last block
#pragma unroll 4
        for (uintptr_t blk = A0 + 16 * (uintptr_t)threadIdx.x; blk < A1; blk += 16 * (uintptr_t)blockDim.x) {
            u32x4 v = *(gptr<const u32x4>)to_global((const u32x4 *)blk);
            if (blk < p || blk + 16 > e) {
                // keep bytes [lo, hi) of the block
                const uint32_t lo = blk < p ? (uint32_t)(p - blk) : 0u, hi = blk + 16 > e ? (uint32_t)(e - blk) : 16u;
                uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    uint32_t m = 0xFFFFFFFFu;
                    if (lo > 4u * j) m = lo >= 4u * j + 4 ? 0u : m << (8 * (lo - 4 * j));
                    if (hi < 4u * j + 4) m &= hi <= 4u * j ? 0u : 0xFFFFFFFFu >> (8 * (4 * j + 4 - hi));
                    w[j] &= m;
                }
                v.x = w[0];
                v.y = w[1];
                v.z = w[2];
                v.w = w[3];
            }
            acc = mul_x4k(mx, acc) ^ crc16b(t8, v.x, v.y, v.z, v.w);
            my_end = blk + 16;
        }
        if (my_end) acc = crc_shift(acc, (uint64_t)(A1 - my_end));
        if (acc) atomicXor(&red, acc);
        __syncthreads();
        if (threadIdx.x == 0) seg_crc[s] = crc_shift(crc_unshift(red, (uint32_t)(A1 - e)), seg_tail[s]);
        __syncthreads();
    }
}

// Per member m: segments [mseg[m], mseg[m+1]) and, when tile_crc is given,
// resolve tiles [mtile[2m], mtile[2m+1]) (their CRCs come from
// k_resolve_body, shifted the same way), total bytes mlen[m].  One workgroup
// per member.
__global__ void __launch_bounds__(256) k_crc_members(const uint32_t *__restrict__ seg_crc,
                                                     const uint32_t *__restrict__ mseg, const uint64_t *__restrict__ mlen,
                                                     uint32_t nmem, uint32_t *__restrict__ crc_out,
                                                     const uint32_t *__restrict__ tile_crc = nullptr,
                                                     const uint32_t *__restrict__ mtile = nullptr) {
    __shared__ uint32_t red;
    for (uint32_t m = blockIdx.x; m < nmem; m += gridDim.x) {
        if (threadIdx.x == 0) red = 0;
        __syncthreads();
        uint32_t c = 0;
        for (uint32_t s = mseg[m] + threadIdx.x; s < mseg[m + 1]; s += blockDim.x) c ^= seg_crc[s];
        if (tile_crc)
            for (uint32_t s = mtile[2 * m] + threadIdx.x; s < mtile[2 * m + 1]; s += blockDim.x) c ^= tile_crc[s];
        if (c) atomicXor(&red, c);
        __syncthreads();
        if (threadIdx.x == 0) crc_out[m] = ~(crc_shift(0xFFFFFFFFu, mlen[m]) ^ red);
        __syncthreads();
    }
}

}  // namespace gz
