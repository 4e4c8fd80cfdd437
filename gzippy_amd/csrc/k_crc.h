// k_crc.h — gzip CRC32 (reflected 0xEDB88320) on the device.
//
// Stage 1 (k_crc_segments): every workgroup folds one segment of <= 64 KiB;
// each of its 256 lanes runs a slicing-by-8 table CRC over 256 contiguous
// bytes from a zero register.  The lane CRCs combine with the GF(2) shift
// crc(A||B) = crc(A) * x^(8|B|) mod P ^ crc(B) (combine_crc32,
// crc32.rs:127-133; polynomial_multiply_modulo :64; x_power_modulo :104):
// lane l's CRC is multiplied by the constant x^(8 * 256 * (255 - l)) and the
// 256 products are XORed.  A segment shorter than 64 KiB is placed at the end
// of the 64 KiB frame behind zeros, which a raw CRC register ignores.
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

// c_kseg[l] = x^(8 * 256 * (255 - l)) mod P: lane l's shift to the end of a
// full 64 KiB segment, filled by the host
__constant__ uint32_t c_kseg[256];

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

// Segments: seg_off[i] bytes into `base`, seg_len[i] <= CRC_SEG.  Every
// segment is folded as the last seg_len bytes of a 64 KiB frame whose first
// bytes are zero: leading zeros leave a raw CRC register at 0, so the frame's
// CRC is the segment's.  The frame is staged in LDS (coalesced loads), then
// lane l folds frame bytes [256 l, 256 l + 256) with slicing-by-8 and the
// lanes combine with the constant shifts c_kseg.  A full, 16-byte-aligned
// segment is staged with 16-byte loads; any other with dword loads realigned
// by v_alignbit (only dwords that hold segment bytes are read).
__global__ void __launch_bounds__(256) k_crc_segments(const uint8_t *__restrict__ base, const uint64_t *__restrict__ seg_off,
                                                      const uint32_t *__restrict__ seg_len,
                                                      const uint64_t *__restrict__ seg_tail, uint32_t nseg,
                                                      const uint32_t *__restrict__ tables, uint32_t *__restrict__ seg_crc) {
    __shared__ uint32_t t[8][256];
    __shared__ uint32_t buf[256 * 65];  // frame row r (bytes [256 r, 256 r + 256)) at dword 65 r (conflict-free reads)
    __shared__ uint32_t red;
    for (uint32_t i = threadIdx.x; i < 8 * 256; i += blockDim.x) t[i >> 8][i & 255] = tables[i];
    __syncthreads();
    for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        const uint8_t *p = base + seg_off[s];
        const uint32_t len = seg_len[s];
        uint32_t f0 = CRC_SEG - len;  // frame byte of the segment's first byte
        // A segment whose 16-byte chunks fit the frame (lead bytes before it
        // in its first chunk, pad zero bytes after it to the next 16-byte
        // boundary) is staged with 16-byte loads, masked at both ends; the
        // frame then holds the segment followed by `pad` zeros, whose shift
        // is undone at the end.
        const uint32_t lead = (uint32_t)((uintptr_t)p & 15), pad = (uint32_t)(-(uintptr_t)(p + len) & 15);
        const bool chunked = len && lead + len + pad <= CRC_SEG && !(len == CRC_SEG && lead == 0);
        if (chunked) f0 = CRC_SEG - len - pad;
        if (chunked) {
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            const uint32_t fc = f0 - lead;  // frame byte of the first chunk (16-byte aligned)
            const uint32_t nch = (lead + len + pad) >> 4;
            gptr<const u32x4> src = (gptr<const u32x4>)to_global(p - lead);
            for (uint32_t k = threadIdx.x; k < nch; k += blockDim.x) {
                u32x4 v = src[k];
                if (k == 0 || k == nch - 1) {
                    // zero the bytes outside [lead, lead + len) of the chunk run
                    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                    for (int j = 0; j < 4; j++)
#pragma unroll
                        for (int b = 0; b < 4; b++) {
                            const uint32_t q = 16 * k + 4 * j + b;  // byte of the chunk run
                            if (q < lead || q >= lead + len) w[j] &= ~(0xFFu << (8 * b));
                        }
                    v.x = w[0];
                    v.y = w[1];
                    v.z = w[2];
                    v.w = w[3];
                }
                const uint32_t o = fc + 16 * k;
                uint32_t *d = &buf[65 * (o >> 8) + ((o & 255) >> 2)];
                d[0] = v.x;
                d[1] = v.y;
                d[2] = v.z;
                d[3] = v.w;
            }
            // the first data row's bytes before the chunk run read as zero
            const uint32_t r0 = fc >> 8;
            for (uint32_t D = r0 * 64 + threadIdx.x; D < (fc >> 2); D += blockDim.x) buf[65 * (D >> 6) + (D & 63)] = 0;
        } else if (len == CRC_SEG && (((uintptr_t)p) & 15) == 0) {
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            gptr<const u32x4> src = (gptr<const u32x4>)to_global(p);
#pragma unroll 4
            for (uint32_t k = 0; k < 16; k++) {
                const uint32_t o = k * 4096 + threadIdx.x * 16;  // byte offset in the segment
                const u32x4 v = src[o >> 4];
                uint32_t *d = &buf[65 * (o >> 8) + ((o & 255) >> 2)];
                d[0] = v.x;
                d[1] = v.y;
                d[2] = v.z;
                d[3] = v.w;
            }
        } else if (len) {
            // frame dword D = bytes [4D, 4D + 4) = segment bytes [4D - f0, ..):
            // aligned source dwords a and a + 1, shifted by the misalignment
            const uintptr_t g0 = (uintptr_t)p - f0;  // address of frame byte 0 (never dereferenced)
            const uint32_t sh = (uint32_t)(g0 & 3) * 8;
            const uintptr_t lo_dw = (uintptr_t)p >> 2, hi_dw = ((uintptr_t)p + len - 1) >> 2;  // dwords holding segment bytes
            for (uint32_t D = (f0 >> 2) + threadIdx.x; D < CRC_SEG / 4; D += blockDim.x) {
                const uintptr_t a = (g0 >> 2) + D;
                const uint32_t w0 = a >= lo_dw && a <= hi_dw ? *(gptr<const uint32_t>)to_global((const uint32_t *)(a << 2)) : 0u;
                const uint32_t w1 = sh && a + 1 >= lo_dw && a + 1 <= hi_dw
                                        ? *(gptr<const uint32_t>)to_global((const uint32_t *)((a + 1) << 2))
                                        : 0u;
                uint32_t v = sh ? __builtin_amdgcn_alignbit(w1, w0, sh) : w0;
                if (4 * D < f0) v &= 0xFFFFFFFFu << (8 * (f0 - 4 * D));  // bytes before the segment
                buf[65 * (D >> 6) + (D & 63)] = v;
            }
            // the first data row's bytes before the segment read as zero
            const uint32_t r0 = f0 >> 8;
            for (uint32_t D = r0 * 64 + threadIdx.x; D < (f0 >> 2); D += blockDim.x) buf[65 * (D >> 6) + (D & 63)] = 0;
        }
        if (threadIdx.x == 0) red = 0;
        __syncthreads();
        uint32_t c = 0;
        if (len && threadIdx.x >= (f0 >> 8)) {  // rows before the segment are all zero: CRC 0
            const uint32_t *my = &buf[65 * threadIdx.x];
#pragma unroll 4
            for (uint32_t j = 0; j < 32; j++) {
                const uint32_t lo = c ^ my[2 * j], hi = my[2 * j + 1];
                c = t[7][lo & 0xFF] ^ t[6][(lo >> 8) & 0xFF] ^ t[5][(lo >> 16) & 0xFF] ^ t[4][lo >> 24] ^
                    t[3][hi & 0xFF] ^ t[2][(hi >> 8) & 0xFF] ^ t[1][(hi >> 16) & 0xFF] ^ t[0][hi >> 24];
            }
            c = gf2_mulmod(c_kseg[threadIdx.x], c);
        }
        if (c) atomicXor(&red, c);
        __syncthreads();
        if (threadIdx.x == 0) seg_crc[s] = crc_shift(chunked ? crc_unshift(red, pad) : red, seg_tail[s]);
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
