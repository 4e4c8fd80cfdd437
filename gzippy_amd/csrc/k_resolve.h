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
// k_chain walks a chain of marker pieces in order, keeping the rolling 32 KiB
// window in LDS, and writes each piece's last 32 KiB (the end window its
// successor needs: get_last_window, chunk_data.rs:1295).  A chain starts at
// the member's first marker piece and at every piece that flipped to u8
// entries (its last 32 KiB are literal, so its end window depends on nothing
// before it); chains are independent workgroups.  k_resolve_body then
// resolves every other entry in parallel, reading windows from the now final
// output.
#pragma once
#include "gz_common.h"

namespace gz {

struct ResolvePiece {
    const uint8_t *slab;
    uint8_t *mout;      // member output base
    uint64_t out_pos;   // piece offset in the member output
    uint64_t n16, n8, u8_off;
};

struct ResolveTile {
    uint32_t piece;
    uint32_t pad;
    uint64_t t0, t1;
};

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
        const uint32_t p0 = chains[c], p1 = chains[c + 1];
        {
            const ResolvePiece f = rp[p0];
            // a flipped first piece (u8 entries present, >= 32768 entries)
            // never reads the window before it: leave it unloaded, since that
            // output may still be being written by another chain
            if (!(f.u8_off != 0 && f.n16 + f.n8 >= WIN))
            for (uint32_t j = threadIdx.x; j < WIN; j += blockDim.x) {
                int64_t q = (int64_t)f.out_pos - (int64_t)WIN + j;
                win[0][j] = q >= 0 ? to_global(f.mout)[q] : 0;
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

__global__ void __launch_bounds__(256) k_resolve_body(const ResolvePiece *__restrict__ rp,
                                                      const ResolveTile *__restrict__ tiles, uint32_t ntiles) {
    for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
        const ResolveTile tl = tiles[b];
        const ResolvePiece r = rp[tl.piece];
        gptr<const uint8_t> wbase = to_global((const uint8_t *)r.mout) + r.out_pos;  // window = wbase[-WIN .. 0)
        const bool near_start = r.out_pos < WIN;
        gptr<const uint16_t> e16 = (gptr<const uint16_t>)to_global(r.slab);
        gptr<const uint8_t> s8 = to_global(r.slab);
        gptr<uint8_t> mout = to_global(r.mout);
        for (uint64_t t = tl.t0 + threadIdx.x; t < tl.t1; t += blockDim.x) {
            uint8_t v;
            if (t < r.n16) {
                uint32_t e = e16[t];
                if (e < 256) v = (uint8_t)e;
                else {
                    int64_t q = (int64_t)(e - MARKER_BASE) - (int64_t)WIN;
                    v = (near_start && (int64_t)r.out_pos + q < 0) ? 0 : wbase[q];
                }
            } else {
                v = s8[r.u8_off + (t - r.n16)];
            }
            mout[r.out_pos + t] = v;
        }
    }
}

}  // namespace gz
