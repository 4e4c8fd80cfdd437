// graph.h — a batch of members as one replayable hipGraph.
//
// Included at the end of gzgpu.hip (same translation unit).
//
// gzgpu_batch_graph_create fixes a layout: where each member's compressed
// bytes lie in d_in and where its output goes in d_out.  It captures, on the
// context's stream, every step of decoding that batch with no host decision
// in between.  A member whose decode cost is more than one resident
// wavefront's share of the batch is split into pieces, planned on the device:
//   1. k_graph_clear, k_graph_prep: queues, counters and planning cells are
//      cleared by a kernel (no memset node); one thread per member parses
//      the gzip header from device memory (gzip_format.rs:47), reads the
//      trailer and writes the member's first piece (DIRECT, from the deflate
//      start, into the member's output);
//   2. split members only: the block finder over their compressed bytes
//      (k_bf_stage_a / _b, k_bf_estimate; the ranges are fixed at creation);
//   3. k_graph_hash_hits, k_graph_bucket: every hit goes to its member's
//      grid cell (np cells of equal compressed bits); a cell keeps its first
//      dynamic header, its first stored block that chains (the block after
//      its payload is another hit, found in a device-side hash set), its last
//      real header, and the sum of its hits' output estimates;
//   4. k_graph_plan, one workgroup per split member: marker piece c starts
//      at cell c's first hit (a cell without one merges into the piece
//      before it), the piece before it stops there, and its marker slab is
//      sized from the cells' estimates scaled to the trailer's ISIZE;
//   5. k_inflate<uint16_t>: one persistent launch over every piece, pulled
//      longest first from the device-side queue;
//   6. k_graph_link_local, one workgroup per split member: piece c is
//      accepted when the piece before it stopped inside c's accepted start
//      range (try_speculative_decode_candidate, chunk_fetcher.rs:2506); a
//      piece that is not, whose predecessor was, is decoded again in place
//      from the predecessor's end over its own range, in a second
//      k_inflate<uint16_t> launch (its piece count written on the device);
//   7. k_graph_link_final: the link check again.  At the first piece still
//      not accepted, or that overflowed its slab, the member's remainder
//      becomes ONE re-decode piece from the last confirmed boundary (the end
//      of the last accepted piece, or the overflowed piece's own accepted
//      start) with a slab for the whole member, and the pieces after it are
//      dropped; a third launch decodes those.  That piece reaches the final
//      block or a real error, so a valid stream always finishes;
//   8. k_graph_resolve_prep, k_chain_scan_*, k_resolve_body: each split
//      member's pieces in stream order (dropped ones empty) form one window
//      chain, resolved as a parallel prefix (get_last_window,
//      chunk_data.rs:1295), then every piece's body is resolved
//      (replace_markers.rs:129);
//   9. k_graph_members, k_graph_seglen, k_crc_segments, k_crc_members,
//      k_graph_verify: CRC32 per member and the trailer checks.
// gzgpu_batch_graph_launch replays it with one hipGraphLaunch and copies the
// per-member status, length and CRC back.  The caller may rewrite the
// contents of d_in between launches; offsets, lengths and capacities stay.
// Piece counts come from the contents at creation (the trailers' ISIZE): a
// schedule only, so any contents decode exactly, other contents perhaps less
// evenly balanced.

namespace gz {

constexpr uint32_t GRAPH_NO_SPLIT = 0xFFFFFFFFu;
// counters cleared on every replay
enum : uint32_t {
    GC_QUEUE1 = 0, GC_QUEUE2, GC_QUEUE3,  // inflate queues of the three launches
    GC_SURV, GC_HITS,                     // block finder
    GC_PIECES, GC_REDO,                   // statistics: marker pieces planned, re-decodes
    GC_N2, GC_N3,                         // pieces of the second and third inflate launch
    GC_N = 16
};

struct GraphMember {
    uint64_t in_off, in_len;    // bytes, relative to the aligned input base
    uint64_t out_off, out_cap;  // bytes, relative to d_out
    uint32_t split;             // index into the split table, or GRAPH_NO_SPLIT
    uint32_t pad;
};

struct GraphHead {
    uint32_t status;  // GZGPU_OK or the header error
    uint32_t exp_crc, exp_isize, pad;
};

// A split member, fixed at creation.  Its pieces: the member's own index
// (piece 0, DIRECT), marker pieces first .. first + np - 2 (cells 1 .. np-1)
// and the re-decode piece `redo`.  Chain slots slot0 .. slot0 + np - 1 hold
// marker pieces 1 .. np - 1, then the re-decode piece; sequence positions
// pos0 .. pos0 + np add piece 0 in front.
struct GraphSplit {
    uint32_t member, np, first, redo, slot0, cell0, pos0, pad;
    uint64_t slab_off, slab_bytes;  // marker slabs of cells 1 .. np-1 in the slab pool
    uint64_t redo_off, redo_bytes;  // two re-decode slabs of redo_bytes each (an overflowed piece's, the remainder's)
};

// Per split member, written on the device every replay.
struct GraphSplitRun {
    uint64_t ds, de, limit, spacing;  // deflate bits [ds, de) (de: trailer start), readable bits, grid spacing
    uint32_t ok, isize;               // header parsed; trailer ISIZE
    uint32_t fail, redo;              // decode status (GZGPU_*; GZGPU_PENDING while a re-decode runs); re-decoded
    uint32_t first_bad, why;          // link check: first piece not accepted (np: none); 1 unlinked, 2 overflow, 3 error
};

// Planning cell: its first dynamic-header hit that may start a piece
// (bit_offset << 27 | hit index; ~0 none), its last real block header
// (dynamic, or a chained stored block; same form, 0 none), its first
// final-block header (bit; ~0 none), its first chained stored-block hit, the
// sum (8.8 fixed point) and count of its hits' output estimates.
struct GraphCell {
    uint64_t key, last, fin;
    uint64_t skey;  // its first chained stored hit (same form as key; ~0 none)
    uint64_t send;  // the furthest payload end (bit) of the chained stored hits in it (0 none)
    uint32_t rsum, rcnt;
};

// Marker piece of cell c: accepted start range (a start inside a block: the
// grid point, its sync point known after the launch), the header of the
// block it starts inside, where it stops (and the header of the block its
// successor starts inside, ~0 none), the cumulative output estimate of cells
// 0 .. c, the cell where the next piece starts, whether a piece starts in
// the cell, whether it starts inside a block and whether the link check
// dropped it.
struct GraphAcc {
    uint64_t lo, hi, hdr, stop, stop_hdr;
    double ce;
    uint32_t used, mid, dropped, end;
    uint32_t linked, pred;  // the link check: accepted after the first launch; the piece before it
    uint64_t plan_lo;       // the planned start (diagnostics: a piece re-decoded in place has a new one)
    uint32_t plan_mid, pad;
};

constexpr uint32_t GRAPH_PLAN_THREADS = 256, GRAPH_PREP_THREADS = 1024;

// Inclusive scan of one value per thread over the workgroup (buf: one slot
// per thread).  On return buf holds every thread's inclusive value.
template <typename T, typename Op>
__device__ T block_scan_incl(T v, T *buf, Op op) {
    const uint32_t t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (uint32_t d = 1; d < blockDim.x; d <<= 1) {
        const T x = t >= d ? buf[t - d] : T{};
        __syncthreads();
        if (t >= d) buf[t] = op(buf[t], x);
        __syncthreads();
    }
    return buf[t];
}

__device__ __forceinline__ PieceDesc graph_dummy_piece(uint32_t member) {
    // starts at its limit: the kernel reports a truncated stream at once
    PieceDesc d{};
    d.start_bit = d.limit_bit = 0;
    d.stop_bit = ~0ull;
    d.dst = nullptr;
    d.cap = 0;
    d.flags = PF_DIRECT;
    d.member = member;
    return d;
}

__global__ void k_graph_clear(GraphCell *__restrict__ cells, uint32_t ncells, uint32_t *__restrict__ counters,
                              uint64_t *__restrict__ htab, uint32_t ht_size) {
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 < GC_N) counters[i0] = 0;
    for (uint32_t i = i0; i < ncells; i += gridDim.x * blockDim.x)
        cells[i] = GraphCell{~0ull, 0ull, ~0ull, ~0ull, 0ull, 0u, 0u};
    for (uint32_t i = i0; i < ht_size; i += gridDim.x * blockDim.x) htab[i] = ~0ull;
}

// ---- the stored-block chain test on the device ---------------------------
// A LEN / ~LEN coincidence in compressed data passes the stored-block test
// about once per 2^16 byte positions, so stored hits are only starts when
// chained: the block after the payload is another hit (a dynamic header at
// exactly that bit, or a stored block whose LEN field follows it) or the
// member ends there -- the host planner's test.  Every hit goes into an
// open-addressing set: a dynamic header by its bit, a stored block by the
// byte of its LEN field.
constexpr uint32_t GRAPH_HT_PROBES = 64;
__device__ __forceinline__ uint64_t graph_hit_key(const gzgpu_boundary &b) {
    return b.kind == 2 ? b.bit_offset << 1 : (((b.seek_bit + 3) >> 3) << 1) | 1u;
}
__device__ __forceinline__ uint32_t graph_hash(uint64_t k) {
    k *= 0x9E3779B97F4A7C15ull;
    return (uint32_t)(k >> 32);
}
__device__ bool graph_ht_has(const uint64_t *__restrict__ htab, uint32_t mask, uint64_t key) {
    uint32_t s = graph_hash(key) & mask;
    for (uint32_t i = 0; i < GRAPH_HT_PROBES; i++, s = (s + 1) & mask) {
        const uint64_t v = htab[s];
        if (v == key) return true;
        if (v == ~0ull) return false;
    }
    return false;
}

__global__ void k_graph_hash_hits(const gzgpu_boundary *__restrict__ hits, const uint32_t *__restrict__ counters,
                                  uint32_t hit_cap, uint64_t *__restrict__ htab, uint32_t mask) {
    const uint32_t counted = counters[GC_HITS];
    const uint32_t nh = counted < hit_cap ? counted : hit_cap;
    for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += gridDim.x * blockDim.x) {
        const uint64_t key = graph_hit_key(hits[h]);
        uint32_t s = graph_hash(key) & mask;
        for (uint32_t i = 0; i < GRAPH_HT_PROBES; i++, s = (s + 1) & mask) {
            const uint64_t prev = atomicCAS((unsigned long long *)&htab[s], ~0ull, (unsigned long long)key);
            if (prev == ~0ull || prev == key) break;  // set full along the probe run: the key is left out
        }
    }
}

__global__ void k_graph_prep(const uint8_t *__restrict__ in, uint8_t *out, const GraphMember *__restrict__ gm,
                             uint32_t n, PieceDesc *__restrict__ pieces, GraphHead *__restrict__ head,
                             const GraphSplit *__restrict__ sp, GraphSplitRun *__restrict__ run) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GraphMember g = gm[i];
    const uint8_t *p = in + g.in_off;
    const uint64_t L = g.in_len;
    PieceDesc d{};
    d.stop_bit = ~0ull;
    d.dst = out + g.out_off;
    d.cap = g.out_cap;
    d.flags = PF_DIRECT;
    d.member = i;
    GraphHead h{};
    h.status = GZGPU_OK;
    // RFC 1952 header: magic, CM = 8, FLG; FEXTRA, FNAME, FCOMMENT, FHCRC
    uint64_t off = 10;
    if (L < 18 || p[0] != 0x1f || p[1] != 0x8b || p[2] != 8) {
        h.status = GZGPU_INVALID_HEADER;
    } else {
        const uint32_t flg = p[3];
        if (flg & 4) {
            if (off + 2 > L) h.status = GZGPU_INVALID_HEADER;
            else off += 2 + ((uint64_t)p[off] | (uint64_t)p[off + 1] << 8);
        }
        for (uint32_t f = 8; f <= 16 && h.status == GZGPU_OK; f <<= 1) {
            if (!(flg & f)) continue;
            while (off < L && p[off]) off++;
            if (off >= L) h.status = GZGPU_INVALID_HEADER;
            else off++;
        }
        if (h.status == GZGPU_OK && (flg & 2)) off += 2;
        if (h.status == GZGPU_OK && off + 8 > L) h.status = GZGPU_INVALID_FORMAT;
    }
    if (h.status == GZGPU_OK) {
        d.start_bit = (g.in_off + off) * 8;
        d.limit_bit = (g.in_off + L) * 8;
        const uint8_t *t = p + L - 8;
        h.exp_crc = (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16 | (uint32_t)t[3] << 24;
        h.exp_isize = (uint32_t)t[4] | (uint32_t)t[5] << 8 | (uint32_t)t[6] << 16 | (uint32_t)t[7] << 24;
    } else {
        // start at the limit: the kernel reports a truncated stream at once
        d.start_bit = d.limit_bit = g.in_off * 8;
        d.cap = 0;
    }
    pieces[i] = d;
    head[i] = h;
    if (g.split != GRAPH_NO_SPLIT) {
        GraphSplitRun r{};
        r.ok = h.status == GZGPU_OK;
        r.ds = d.start_bit;
        r.de = (g.in_off + L - (L >= 8 ? 8 : L)) * 8;
        r.limit = d.limit_bit;
        const uint32_t np = sp[g.split].np;
        r.spacing = r.de > r.ds ? (r.de - r.ds) / np : 0;
        if (r.spacing == 0) r.ok = 0;  // nothing to split: piece 0 decodes the member
        r.isize = h.exp_isize;
        r.fail = GZGPU_OK;
        r.redo = 0;
        r.first_bad = np;
        r.why = 0;
        run[g.split] = r;
    }
}

// One thread per block-finder hit: the split member whose finder range holds
// it (ranges sorted by start, two per member), then its grid cell.
__global__ void k_graph_bucket(const gzgpu_boundary *__restrict__ hits, const uint32_t *__restrict__ counters,
                               uint32_t hit_cap, const uint64_t *__restrict__ sp_lo, const uint64_t *__restrict__ sp_hi,
                               const uint32_t *__restrict__ sp_idx, uint32_t nranges, const GraphSplit *__restrict__ sp,
                               const GraphSplitRun *__restrict__ run, GraphCell *__restrict__ cells,
                               const uint64_t *__restrict__ htab, uint32_t ht_mask) {
    const uint32_t counted = counters[GC_HITS];
    const uint32_t nh = counted < hit_cap ? counted : hit_cap;
    for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += gridDim.x * blockDim.x) {
        const gzgpu_boundary b = hits[h];
        uint32_t lo = 0, hi = nranges;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (sp_lo[mid] <= b.bit_offset) lo = mid;
            else hi = mid;
        }
        if (b.bit_offset < sp_lo[lo] || b.bit_offset >= sp_hi[lo]) continue;
        const uint32_t s = sp_idx[lo];
        const GraphSplitRun R = run[s];
        if (!R.ok || b.bit_offset <= R.ds || b.seek_bit >= R.de) continue;
        float ratio;
        uint64_t pay_end = 0;
        if (b.kind == 2) {
            ratio = exp2f((float)(b.info >> 20) * (1.0f / 256.0f) - 8.0f);
        } else if (b.kind == 0 && (b.info & 1u)) {
            // chained: the next block starts where the payload ends
            const uint64_t next = (((b.seek_bit + 3) >> 3) + 4 + (b.info >> 16)) << 3;
            const bool chained = next == R.de || graph_ht_has(htab, ht_mask, next << 1) ||
                                 graph_ht_has(htab, ht_mask, ((((next >> 3) + 1)) << 1) | 1u);
            if (!chained) continue;
            ratio = 0.125f;  // a byte per 8 bits
            pay_end = next;
        } else {
            continue;
        }
        const bool final_hdr = b.kind == 2 && ((b.info >> HIT_FINAL_SHIFT) & 1u);
        const uint32_t np = sp[s].np;
        uint64_t c = (b.bit_offset - R.ds) / R.spacing;
        if (c >= np) c = np - 1;
        GraphCell *cell = cells + sp[s].cell0 + c;
        const float fx = ratio * 256.0f;
        const unsigned long long key = (unsigned long long)(b.bit_offset << 27 | h);
        atomicAdd(&cell->rsum, fx < 65535.0f ? (uint32_t)fx : 65535u);
        atomicAdd(&cell->rcnt, 1u);
        atomicMax((unsigned long long *)&cell->last, key);
        if (pay_end) atomicMax((unsigned long long *)&cell->send, (unsigned long long)pay_end);
        // a final block's header is never a piece start (the stop rule stops
        // at non-final headers only), only the header of starts inside it
        if (final_hdr) atomicMin((unsigned long long *)&cell->fin, (unsigned long long)b.bit_offset);
        else if (c >= 1) atomicMin((unsigned long long *)(b.kind == 2 ? &cell->key : &cell->skey), key);
    }
}

__device__ __forceinline__ uint64_t graph_slab_bytes(uint64_t entries) {
    return ((2 * (entries + 65536) + 512) + 255) & ~255ull;  // marker_slab_bytes
}

// Starts inside a block: at least MID_HEAD bits after the block's header
// (a dynamic header is rarely longer) and MID_TAIL bits before the next
// header; the same rule as the host planner.
constexpr uint64_t GRAPH_MID_HEAD = 2048, GRAPH_MID_TAIL = 2048;
// Marker slab per piece: its cells' output estimate x GRAPH_SLAB_FACTOR +
// 4096 entries (a piece whose successor's start fails runs on to the next
// block; an overflow costs a re-decode of the member's remainder)
constexpr double GRAPH_SLAB_FACTOR = 2.0;

// One workgroup per split member: piece starts, stops and marker slabs.
// Cell c (c >= 1) starts a piece at its first hit when that lies within an
// eighth of a spacing of the grid point; else inside the block that holds
// the grid point (a MIDBLOCK piece: it reads the header of that block, the
// member's last real hit before the point, and starts at the sync point of
// the grid point); else at its first hit anywhere; else nowhere (the piece
// before runs on).
__global__ void __launch_bounds__(GRAPH_PLAN_THREADS)
    k_graph_plan(const GraphSplit *__restrict__ sp, const GraphSplitRun *__restrict__ run,
                 const GraphCell *__restrict__ cells, GraphAcc *__restrict__ acc,
                 const gzgpu_boundary *__restrict__ hits, const GraphMember *__restrict__ gm,
                 PieceDesc *__restrict__ pieces, uint8_t *__restrict__ pool, uint32_t *__restrict__ counters,
                 uint32_t miss_every) {
    __shared__ uint64_t sbuf[GRAPH_PLAN_THREADS];
    __shared__ double dbuf[GRAPH_PLAN_THREADS];
    __shared__ uint64_t stop0, stop0_hdr;
    const uint32_t t = threadIdx.x, NT = blockDim.x;
    const GraphSplit S = sp[blockIdx.x];
    const GraphSplitRun R = run[blockIdx.x];
    const uint32_t np = S.np;
    GraphAcc *A = acc + S.cell0;
    const GraphCell *C = cells + S.cell0;
    auto add = [](uint64_t a, uint64_t b) { return a + b; };
    auto mx = [](uint64_t a, uint64_t b) { return a > b ? a : b; };
    if (!R.ok) {
        for (uint32_t c = 1 + t; c < np; c += NT) {
            pieces[S.first + c - 1] = graph_dummy_piece(S.member);
            A[c].used = 0;
            A[c].dropped = 0;
        }
        return;
    }
    if (t == 0) {
        stop0 = ~0ull;
        stop0_hdr = ~0ull;
    }
    const bool midblock = miss_every == 0;  // the miss hook shifts hit starts only
    // pass 1: where each cell's piece starts (the last real header before the
    // cell: an exclusive max-scan of the cells' last headers), and the k-th
    // start (the miss_every test hook plans every n-th start one bit late)
    uint64_t lcarry = 0, ucarry = 0, scarry = 0;
    for (uint32_t base = 0; base < np; base += NT) {
        const uint32_t c = base + t;
        block_scan_incl<uint64_t>(c < np ? C[c].last : 0ull, sbuf, mx);
        const uint64_t lprev = mx(lcarry, t ? sbuf[t - 1] : 0ull);
        const uint64_t lchunk = sbuf[NT - 1];
        __syncthreads();
        // the furthest end of a chained stored payload that starts before the cell
        block_scan_incl<uint64_t>(c < np ? C[c].send : 0ull, sbuf, mx);
        const uint64_t sprev = mx(scarry, t ? sbuf[t - 1] : 0ull);
        const uint64_t schunk = sbuf[NT - 1];
        __syncthreads();
        GraphAcc a{};
        a.end = np;
        a.stop = ~0ull;
        a.stop_hdr = ~0ull;
        if (c >= 1 && c < np) {
            const uint64_t g = R.ds + (uint64_t)c * R.spacing;
            uint64_t dkey = C[c].key;
            const uint64_t skey = C[c].skey;
            uint64_t hd = dkey != ~0ull ? dkey >> 27 : ~0ull;       // first dynamic header
            const uint64_t hs = skey != ~0ull ? skey >> 27 : ~0ull;  // first chained stored block
            // a dynamic header inside a chained stored block's payload is
            // data (a gzip of deflate data), not a block of this stream:
            // shadowed when a payload that starts before the cell, or the
            // cell's own stored blocks after the first, reach past it
            if (hd != ~0ull && (hd < sprev || (hs < hd && hd < C[c].send))) {
                hd = ~0ull;
                dkey = ~0ull;
            }
            const uint64_t hb = hd < hs ? hd : hs;
            const uint64_t next = hb < C[c].fin ? hb : C[c].fin;     // first real header after the grid point
            uint64_t hdr = R.ds;
            bool dyn = true;
            if (lprev) {
                const gzgpu_boundary &h = hits[(uint32_t)(lprev & ((1u << 27) - 1))];
                hdr = h.seek_bit;
                dyn = h.kind == 2;
            }
            // a chained stored block near the grid point, else a dynamic
            // header near it, else a start inside the block holding it, else
            // the cell's first chained stored block or dynamic header
            uint64_t key = ~0ull;
            if (hs != ~0ull && hs - g <= R.spacing / 8) key = skey;
            else if (hd != ~0ull && hd - g <= R.spacing / 8) key = dkey;
            if (key != ~0ull) {
                a.used = 1;
            } else if (midblock && dyn && g >= hdr + GRAPH_MID_HEAD && (next == ~0ull || next >= g + GRAPH_MID_TAIL)) {
                a.used = 1;
                a.mid = 1;
                a.lo = a.hi = g;
                a.hdr = hdr;
            } else if (hb != ~0ull) {
                a.used = 1;
                key = hs != ~0ull ? skey : dkey;
            }
            if (a.used && !a.mid) {
                const gzgpu_boundary &h = hits[(uint32_t)(key & ((1u << 27) - 1))];
                a.lo = h.bit_offset;
                a.hi = h.seek_bit;
            }
        }
        const uint64_t k = ucarry + block_scan_incl<uint64_t>(a.used ? 1ull : 0ull, sbuf, add);
        const uint64_t uchunk = sbuf[NT - 1];
        __syncthreads();
        if (a.used && !a.mid && miss_every && k % miss_every == 0) a.lo = a.hi = a.hi + 1;
        a.plan_lo = a.lo;
        a.plan_mid = a.mid;
        if (c < np) A[c] = a;
        lcarry = mx(lcarry, lchunk);
        scarry = mx(scarry, schunk);
        ucarry += uchunk;
    }
    if (t == 0) atomicAdd(&counters[GC_PIECES], (uint32_t)ucarry);
    __syncthreads();
    // pass 2: every start tells the piece before it where to stop (and, for
    // a start inside a block, that block's header: STOPMID)
    uint64_t carry = 0;
    for (uint32_t base = 0; base < np; base += NT) {
        const uint32_t c = base + t;
        const bool used = c < np && A[c].used;
        block_scan_incl<uint64_t>(used ? (uint64_t)c : 0ull, sbuf, mx);
        const uint64_t prev_in_chunk = t ? sbuf[t - 1] : 0ull;
        const uint64_t chunk = sbuf[NT - 1];
        __syncthreads();
        const uint64_t pred = mx(prev_in_chunk, carry);  // last start before c (0: piece 0)
        if (used) {
            const GraphAcc a = A[c];
            const uint64_t sh = a.mid ? a.hdr : ~0ull;
            if (pred == 0) {
                stop0 = a.lo;
                stop0_hdr = sh;
            } else {
                A[pred].stop = a.lo;
                A[pred].stop_hdr = sh;
                A[pred].end = c;
            }
        }
        carry = mx(carry, chunk);
        __syncthreads();
    }
    // pass 3: output estimate per cell (its hits' mean entries per bit, else
    // the last cell before it with hits, else the member's mean), running sum
    const double mean = (double)R.isize / (double)(R.de - R.ds);
    double dcarry = 0;
    uint64_t kcarry = 0;  // last cell with hits, + 1
    for (uint32_t base = 0; base < np; base += NT) {
        const uint32_t c = base + t;
        const bool known = c < np && C[c].rcnt != 0;
        uint64_t kn = block_scan_incl<uint64_t>(known ? (uint64_t)c + 1 : 0ull, sbuf, mx);
        const uint64_t kchunk = sbuf[NT - 1];
        __syncthreads();
        kn = mx(kn, kcarry);
        double e = 0;
        if (c < np) {
            const double r = kn ? (double)C[kn - 1].rsum / (256.0 * (double)C[kn - 1].rcnt) : mean;
            const uint64_t cb = c + 1 < np ? R.spacing : (R.de - R.ds) - (uint64_t)(np - 1) * R.spacing;
            e = r * (double)cb;
        }
        const double ce = dcarry + block_scan_incl<double>(e, dbuf, [](double a, double b) { return a + b; });
        const double dchunk = dbuf[NT - 1];
        __syncthreads();
        if (c < np) A[c].ce = ce;
        dcarry += dchunk;
        kcarry = mx(kcarry, kchunk);
    }
    __syncthreads();
    const double total = dcarry;
    const uint64_t cap = gm[S.member].out_cap;
    const double want = (double)(R.isize < cap ? (uint64_t)R.isize : cap);
    const double scale = total > 0 ? want / total : 0.0;
    if (t == 0) {
        PieceDesc &p0 = pieces[S.member];
        p0.stop_bit = stop0;
        if (stop0_hdr != ~0ull) {
            p0.flags |= PF_STOPMID;
            p0.stop_hdr = stop0_hdr;
        }
    }
    // pass 4: marker slabs and descriptors
    uint64_t ocarry = 0;
    for (uint32_t base = 0; base < np; base += NT) {
        const uint32_t c = base + t;
        const bool used = c < np && A[c].used;
        uint64_t bytes = 0;
        if (used) {
            const GraphAcc a = A[c];
            const double est = (A[a.end - 1].ce - A[c - 1].ce) * scale;
            bytes = graph_slab_bytes((est > 0 ? (uint64_t)(GRAPH_SLAB_FACTOR * est) : 0) + 4096);
        }
        const uint64_t incl = block_scan_incl<uint64_t>(bytes, sbuf, add);
        const uint64_t chunk = sbuf[NT - 1];
        __syncthreads();
        if (c >= 1 && c < np) {
            if (used) {
                const GraphAcc a = A[c];
                const uint64_t off = ocarry + incl - bytes;
                PieceDesc d{};
                d.start_bit = a.mid ? a.lo : a.hi;
                d.stop_bit = a.stop;
                d.limit_bit = R.limit;
                d.dst = pool + S.slab_off + off;
                d.cap = off + bytes <= S.slab_bytes ? bytes : 0;  // no room: overflow, then the re-decode
                d.window = nullptr;
                d.flags = PF_MARKER | (a.mid ? PF_MIDBLOCK : 0u) | (a.stop_hdr != ~0ull ? PF_STOPMID : 0u);
                d.member = S.member;
                d.hdr_bit = a.mid ? a.hdr : 0;
                d.stop_hdr = a.stop_hdr != ~0ull ? a.stop_hdr : 0;
                pieces[S.first + c - 1] = d;
            } else {
                pieces[S.first + c - 1] = graph_dummy_piece(S.member);
            }
        }
        ocarry += chunk;
    }
}

__device__ __forceinline__ bool graph_piece_done(uint32_t st) { return st == PS_STOPPED || st == PS_FINAL; }

// The accepted predecessor end of a piece: its planned range, or for a start
// inside a block the sync point it found (none: nothing links to it).
__device__ __forceinline__ void graph_accepted(const GraphAcc &a, const PieceOut &me, uint64_t &lo, uint64_t &hi) {
    if (a.mid) {
        const bool synced = !(me.status == PS_ERROR && me.err == E_SYNC);
        lo = synced ? me.start_bit : 1;
        hi = synced ? me.start_bit : 0;
    } else {
        lo = a.lo;
        hi = a.hi;
    }
}

// One workgroup per split member, after the first inflate launch: the link
// check of every piece against the one before it.  A piece that is not
// accepted, when the piece before it was and stopped at a block boundary, is
// decoded again in place from that boundary over its own planned range (its
// slab covers it: the boundary lies at or after its planned start), in the
// second launch; every such miss of the member at once, as the host
// pipeline's re-decode round does.
__global__ void __launch_bounds__(GRAPH_PLAN_THREADS)
    k_graph_link_local(const GraphSplit *__restrict__ sp, const GraphSplitRun *__restrict__ run,
                       GraphAcc *__restrict__ acc, const PieceOut *__restrict__ pout, PieceDesc *__restrict__ pieces,
                       uint8_t *__restrict__ pool, uint32_t *__restrict__ counters, uint32_t *__restrict__ order2) {
    __shared__ uint64_t sbuf[GRAPH_PLAN_THREADS];
    __shared__ uint32_t ovf;
    const uint32_t t = threadIdx.x, NT = blockDim.x;
    const GraphSplit S = sp[blockIdx.x];
    if (!run[blockIdx.x].ok) return;
    const uint32_t np = S.np;
    GraphAcc *A = acc + S.cell0;
    auto out_of = [&](uint64_t k) -> const PieceOut & { return k == 0 ? pout[S.member] : pout[S.first + k - 1]; };
    if (t == 0) ovf = np;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < np; base += NT) {
        const uint32_t c = base + t;
        const bool used = c < np && A[c].used;
        block_scan_incl<uint64_t>(used ? (uint64_t)c : 0ull, sbuf, [](uint64_t a, uint64_t b) { return a > b ? a : b; });
        const uint64_t prev_in_chunk = t ? sbuf[t - 1] : 0ull;
        const uint64_t chunk = sbuf[NT - 1];
        __syncthreads();
        const uint64_t pred = prev_in_chunk > carry ? prev_in_chunk : carry;
        if (used) {
            const PieceOut &rp = out_of(pred);
            uint64_t lo, hi;
            graph_accepted(A[c], out_of(c), lo, hi);
            const bool linked = rp.status == PS_STOPPED && rp.end_bit >= lo && rp.end_bit <= hi;
            A[c].linked = linked;
            A[c].pred = (uint32_t)pred;
            if (linked && out_of(c).status == PS_OVERFLOW) atomicMin(&ovf, c);
        }
        if (chunk > carry) carry = chunk;
        __syncthreads();
    }
    // the first piece with an accepted start that overflowed its slab: again
    // in place, from its own start (a start inside a block keeps its header
    // and grid point: the same sync point), into the first half of the
    // member's re-decode slab
    if (t == 0 && ovf < np) {
        const uint32_t pi = S.first + ovf - 1;
        PieceDesc d = pieces[pi];
        d.dst = pool + S.redo_off;
        d.cap = S.redo_bytes;
        pieces[pi] = d;
        order2[atomicAdd(&counters[GC_N2], 1u)] = pi;
        atomicAdd(&counters[GC_REDO], 1u);
    }
    for (uint32_t c = 1 + t; c < np; c += NT) {
        if (!A[c].used || A[c].linked) continue;
        const uint32_t p = A[c].pred;
        const PieceOut &rp = out_of(p);
        if (rp.status != PS_STOPPED || (p != 0 && !A[p].linked)) continue;
        const uint32_t pi = S.first + c - 1;
        PieceDesc d = pieces[pi];
        d.start_bit = rp.end_bit;
        d.flags &= ~PF_MIDBLOCK;
        d.hdr_bit = 0;
        pieces[pi] = d;
        A[c].lo = A[c].hi = rp.end_bit;
        A[c].mid = 0;
        order2[atomicAdd(&counters[GC_N2], 1u)] = pi;
        atomicAdd(&counters[GC_REDO], 1u);
    }
}

// One workgroup per split member, after the second launch: the link check
// again.  At the first piece that is still not accepted, or that did not
// finish (an overflowed slab), the member's remainder becomes ONE re-decode
// piece from the last confirmed boundary, with a slab for the whole member,
// in the third launch; the pieces after it are dropped.  That piece reaches
// the final block or a real error, so a valid stream always finishes.
__global__ void __launch_bounds__(GRAPH_PLAN_THREADS)
    k_graph_link_final(const GraphSplit *__restrict__ sp, GraphSplitRun *__restrict__ run, GraphAcc *__restrict__ acc,
                       const PieceOut *__restrict__ pout, PieceDesc *__restrict__ pieces, uint8_t *__restrict__ pool,
                       uint32_t *__restrict__ counters, uint32_t *__restrict__ order3) {
    __shared__ uint32_t fmin;
    __shared__ uint32_t fdrop;
    const uint32_t t = threadIdx.x, NT = blockDim.x;
    const GraphSplit S = sp[blockIdx.x];
    const uint32_t np = S.np;
    GraphAcc *A = acc + S.cell0;
    auto out_of = [&](uint64_t k) -> const PieceOut & { return k == 0 ? pout[S.member] : pout[S.first + k - 1]; };
    if (!run[blockIdx.x].ok) {
        if (t == 0) {
            const uint32_t st0 = pout[S.member].status;
            run[blockIdx.x].fail = st0 == PS_FINAL ? GZGPU_OK : st0 == PS_OVERFLOW ? GZGPU_OUTPUT_TOO_SMALL
                                                                                   : GZGPU_DECODE_FAILED;
        }
        return;
    }
    if (t == 0) fmin = np;
    __syncthreads();
    if (t == 0 && !graph_piece_done(pout[S.member].status)) atomicMin(&fmin, 0u);
    // the first piece whose start is not confirmed or that did not finish
    for (uint32_t c = 1 + t; c < np; c += NT) {
        if (!A[c].used) continue;
        const PieceOut &rp = out_of(A[c].pred);
        const PieceOut &me = out_of(c);
        uint64_t lo, hi;
        graph_accepted(A[c], me, lo, hi);
        const bool linked = rp.status == PS_STOPPED && rp.end_bit >= lo && rp.end_bit <= hi;
        if (!linked || !graph_piece_done(me.status)) atomicMin(&fmin, c);
    }
    __syncthreads();
    if (t == 0) {
        const uint32_t f = fmin;
        uint32_t status = GZGPU_OK, drop = np, why = 0;
        uint64_t redo_bit = ~0ull;
        bool from_own = false;  // the remainder starts at piece f's own (accepted) start
        if (f == 0) {
            status = pout[S.member].status == PS_OVERFLOW ? GZGPU_OUTPUT_TOO_SMALL : GZGPU_DECODE_FAILED;
            why = pout[S.member].status == PS_OVERFLOW ? 2 : 3;
        } else if (f < np) {
            const PieceOut &rp = out_of(A[f].pred);
            const PieceOut &rc = out_of(f);
            uint64_t lo, hi;
            graph_accepted(A[f], rc, lo, hi);
            drop = f;
            if (!(rp.status == PS_STOPPED && rp.end_bit >= lo && rp.end_bit <= hi)) {
                // the piece before ended elsewhere: at the member's end (the
                // start was inside the final block or past it), or at a true
                // block boundary that this start missed
                if (rp.status != PS_FINAL) redo_bit = rp.end_bit;
                why = 1;
            } else if (rc.status == PS_OVERFLOW) {
                redo_bit = rc.start_bit;  // a confirmed start whose slab was too small
                from_own = true;
                why = 2;
            } else {
                status = GZGPU_DECODE_FAILED;  // an error at a confirmed start: the stream is corrupt
                why = 3;
            }
        }
        fdrop = drop;
        run[blockIdx.x].first_bad = f;
        run[blockIdx.x].why = why;
        if (redo_bit != ~0ull) {
            PieceDesc d{};
            if (from_own) {
                // piece f's own descriptor: a start inside a block must read
                // that block's header and find the same sync point again
                d = pieces[S.first + f - 1];
                d.flags &= ~PF_STOPMID;
                d.stop_hdr = 0;
            } else {
                d.start_bit = redo_bit;  // a block header the piece before stopped at
                d.flags = PF_MARKER;
                d.window = nullptr;
                d.member = S.member;
            }
            d.stop_bit = ~0ull;
            d.limit_bit = run[blockIdx.x].limit;
            d.dst = pool + S.redo_off + S.redo_bytes;  // the second half
            d.cap = S.redo_bytes;
            pieces[S.redo] = d;
            run[blockIdx.x].fail = GZGPU_PENDING;
            run[blockIdx.x].redo = 1;
            order3[atomicAdd(&counters[GC_N3], 1u)] = S.redo;
            atomicAdd(&counters[GC_REDO], 1u);
        } else {
            run[blockIdx.x].fail = status;
        }
    }
    __syncthreads();
    for (uint32_t c = 1 + t; c < np; c += NT) A[c].dropped = A[c].used && c >= fdrop;
}

// One workgroup over every split member: finishes the re-decode check, lays
// the members' pieces out (positions in stream order: piece 0, marker pieces
// 1 .. np-1, the re-decode piece), and writes the chain's ResolvePiece list
// and the body tiles' prefix.
__global__ void __launch_bounds__(GRAPH_PREP_THREADS)
    k_graph_resolve_prep(const GraphSplit *__restrict__ sp, uint32_t nsplit, GraphSplitRun *__restrict__ run,
                         const GraphAcc *__restrict__ acc, const PieceOut *__restrict__ pout,
                         const PieceDesc *__restrict__ pieces, const GraphMember *__restrict__ gm, uint8_t *d_out,
                         const uint32_t *__restrict__ pos_split, uint32_t npos, uint64_t *__restrict__ plen,
                         uint64_t *__restrict__ ppos, ResolvePiece *__restrict__ rp, uint32_t *__restrict__ tile_first,
                         uint32_t nslots, int32_t *__restrict__ mstat, uint64_t *__restrict__ mlen) {
    __shared__ uint64_t sbuf[GRAPH_PREP_THREADS];
    const uint32_t t = threadIdx.x, NT = blockDim.x;
    for (uint32_t s = t; s < nsplit; s += NT) {
        GraphSplitRun &R = run[s];
        if (R.fail == GZGPU_PENDING) {
            const uint32_t st = pout[sp[s].redo].status;
            R.fail = st == PS_FINAL ? GZGPU_OK : st == PS_OVERFLOW ? GZGPU_OUTPUT_TOO_SMALL : GZGPU_DECODE_FAILED;
        }
    }
    __syncthreads();
    auto piece_of = [&](const GraphSplit &S, uint32_t k) { return k == 0 ? S.member : k < S.np ? S.first + k - 1 : S.redo; };
    for (int pass = 0; pass < 2; pass++) {
        // entries of every position (0 for a failed member), and their
        // exclusive prefix over all positions
        uint64_t carry = 0;
        for (uint32_t base = 0; base < npos; base += NT) {
            const uint32_t q = base + t;
            uint64_t L = 0;
            if (q < npos) {
                const uint32_t s = pos_split[q];
                const GraphSplit S = sp[s];
                const GraphSplitRun &R = run[s];
                const uint32_t k = q - S.pos0;
                if (R.fail == GZGPU_OK) {
                    const PieceOut &o = pout[piece_of(S, k)];
                    if (k == 0) L = o.n8;
                    else if (k < S.np) L = acc[S.cell0 + k].used && !acc[S.cell0 + k].dropped ? o.n16 + o.n8 : 0;
                    else L = R.redo ? o.n16 + o.n8 : 0;
                }
                plen[q] = L;
            }
            const uint64_t incl = block_scan_incl<uint64_t>(L, sbuf, [](uint64_t a, uint64_t b) { return a + b; });
            const uint64_t chunk = sbuf[NT - 1];
            __syncthreads();
            if (q < npos) ppos[q] = carry + incl - L;
            carry += chunk;
        }
        __syncthreads();
        // member totals; more than the output holds fails the member
        for (uint32_t s = t; s < nsplit; s += NT) {
            const GraphSplit S = sp[s];
            GraphSplitRun &R = run[s];
            const uint32_t qe = S.pos0 + S.np;  // the re-decode position, the member's last
            const uint64_t T = ppos[qe] + plen[qe] - ppos[S.pos0];
            if (pass == 0 && R.fail == GZGPU_OK && T > gm[S.member].out_cap) R.fail = GZGPU_OUTPUT_TOO_SMALL;
            if (pass == 1) {
                mstat[S.member] = (int32_t)R.fail;
                mlen[S.member] = R.fail == GZGPU_OK ? T : 0;
            }
        }
        __syncthreads();
    }
    // resolve list (one slot per position after piece 0) and body tiles
    uint64_t carry = 0;
    for (uint32_t base = 0; base < npos; base += NT) {
        const uint32_t q = base + t;
        uint64_t nt = 0;
        uint32_t slot = 0;
        bool is_slot = false;
        if (q < npos) {
            const uint32_t s = pos_split[q];
            const GraphSplit S = sp[s];
            const uint32_t k = q - S.pos0;
            if (k >= 1) {
                is_slot = true;
                slot = S.slot0 + k - 1;
                const uint32_t pi = piece_of(S, k);
                const uint64_t L = plen[q];
                const PieceOut &o = pout[pi];
                ResolvePiece r{};
                r.slab = pieces[pi].dst;
                r.mout = d_out + gm[S.member].out_off;
                r.out_pos = ppos[q] - ppos[S.pos0];
                r.n16 = L ? o.n16 : 0;
                r.n8 = L ? o.n8 : 0;
                r.u8_off = L ? o.u8_off : 0;
                const uint32_t qe = S.pos0 + S.np;
                const uint64_t T = ppos[qe] + plen[qe] - ppos[S.pos0];
                r.tail = T - r.out_pos - L;
                r.body = L > WIN ? L - WIN : 0;
                r.win_end = r.mout + r.out_pos;
                r.win_out = 1;
                rp[slot] = r;
                nt = (r.body + RESOLVE_TILE - 1) / RESOLVE_TILE;
            }
        }
        const uint64_t incl = block_scan_incl<uint64_t>(nt, sbuf, [](uint64_t a, uint64_t b) { return a + b; });
        const uint64_t chunk = sbuf[NT - 1];
        __syncthreads();
        if (is_slot) tile_first[slot] = (uint32_t)(carry + incl - nt);
        carry += chunk;
    }
    if (t == 0) tile_first[nslots] = (uint32_t)carry;
}

// Decode status and length of every member that was not split (all of them
// in the plain form).
__global__ void k_graph_members(const GraphMember *__restrict__ gm, uint32_t n, const PieceOut *__restrict__ pout,
                                int32_t *__restrict__ mstat, uint64_t *__restrict__ mlen) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || gm[i].split != GRAPH_NO_SPLIT) return;
    const PieceOut &r = pout[i];
    mstat[i] = r.status == PS_FINAL ? GZGPU_OK : r.status == PS_OVERFLOW ? GZGPU_OUTPUT_TOO_SMALL : GZGPU_DECODE_FAILED;
    mlen[i] = r.status == PS_FINAL ? r.n8 : 0;
}

// Segment s of member seg_member[s] covers its bytes from seg_start[s] to
// the next segment's start (64 KiB for its last), clipped to what the member
// decoded; seg_tail is what the member decoded after it.
__global__ void k_graph_seglen(const uint64_t *__restrict__ mlen, const uint32_t *__restrict__ seg_member,
                               const uint64_t *__restrict__ seg_start, uint32_t nseg, uint32_t *__restrict__ seg_len,
                               uint64_t *__restrict__ seg_tail) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    const uint64_t got = mlen[seg_member[i]];
    const uint64_t s = seg_start[i];
    const uint64_t cap = i + 1 < nseg && seg_member[i + 1] == seg_member[i] ? seg_start[i + 1] - s : CRC_SEG;
    const uint32_t l = got > s ? (uint32_t)(got - s < cap ? got - s : cap) : 0u;
    seg_len[i] = l;
    seg_tail[i] = got > s + l ? got - s - l : 0;
}

__global__ void k_graph_verify(const int32_t *__restrict__ mstat, const uint64_t *__restrict__ mlen,
                               const GraphHead *__restrict__ head, const uint32_t *__restrict__ crc, uint32_t n,
                               int32_t *__restrict__ status, uint64_t *__restrict__ out_len,
                               uint32_t *__restrict__ out_crc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GraphHead h = head[i];
    int32_t st = (int32_t)h.status;
    uint64_t len = 0;
    if (st == GZGPU_OK) {
        st = mstat[i];
        if (st == GZGPU_OK) {
            len = mlen[i];
            if ((uint32_t)len != h.exp_isize) st = GZGPU_SIZE_MISMATCH;
            else if (crc[i] != h.exp_crc) st = GZGPU_CRC_MISMATCH;
        }
    }
    status[i] = st;
    out_len[i] = len;
    out_crc[i] = st == GZGPU_OK || st == GZGPU_CRC_MISMATCH || st == GZGPU_SIZE_MISMATCH ? crc[i] : 0u;
}

}  // namespace gz

struct gzgpu_batch_graph {
    gzgpu_ctx *c = nullptr;
    uint32_t n = 0, nseg = 0;
    // split members: count, pieces of the first launch, chain slots,
    // sequence positions, planning cells, longest chain, finder shape
    uint32_t nsplit = 0, nranges = 0, npieces1 = 0, nslots = 0, npos = 0, ncells = 0, max_chain = 0;
    uint64_t nchunk = 0;
    uint32_t surv_cap = 0, hit_cap = 0, tile_cap = 0, miss_every = 0, ht_size = 0;
    const uint8_t *in_base = nullptr;
    uint8_t *d_out = nullptr;
    uint64_t in_total = 0;
    DevBuf members, pieces, order, pout, counters, head, seg_member, seg_start, seg_off, seg_len, seg_tail, mseg, mlen,
        seg_crc, mcrc, status, olen, ocrc, zero, mstat;
    DevBuf splits, runs, cells, acc, ranges, surv, hits, sp_lo, sp_hi, sp_idx, order2, order3, pos_split, plen, ppos, rps,
        slot_rp, slot_pos, tile_first, tile_crc, scan, pool, htab;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        for (DevBuf *b : {&members, &pieces, &order, &pout, &counters, &head, &seg_member, &seg_start, &seg_off, &seg_len,
                          &seg_tail, &mseg, &mlen, &seg_crc, &mcrc, &status, &olen, &ocrc, &zero, &mstat, &splits, &runs,
                          &cells, &acc, &ranges, &surv, &hits, &sp_lo, &sp_hi, &sp_idx, &order2, &order3, &pos_split, &plen,
                          &ppos, &rps, &slot_rp, &slot_pos, &tile_first, &tile_crc, &scan, &pool, &htab})
            b->release();
    }
};

namespace {

__global__ void k_graph_isize(const uint8_t *__restrict__ in, const GraphMember *__restrict__ gm, uint32_t n,
                              uint32_t *__restrict__ isize) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GraphMember g = gm[i];
    uint32_t v = 0;
    if (g.in_len >= 18) {
        const uint8_t *t = in + g.in_off + g.in_len - 4;
        v = (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16 | (uint32_t)t[3] << 24;
    }
    isize[i] = v;
}

// Buffers of a fixed batch layout, uploaded once.  split == true: members
// costing more than a resident wavefront's share get split pieces (the graph
// form); false: every member whole (the plain form).
void graph_setup(gzgpu_batch_graph *g, const void *d_in, const uint64_t *in_off, const uint64_t *in_len,
                 void *d_out, const uint64_t *out_off, const uint64_t *out_cap, size_t n, bool split) {
    gzgpu_ctx *c = g->c;
    const AlignedIn ai = align_input(d_in);
    g->in_base = ai.base;
    g->d_out = (uint8_t *)d_out;
    uint64_t in_total = 0;
    std::vector<GraphMember> gm(n);
    std::vector<uint32_t> seg_member, mseg;
    std::vector<uint64_t> seg_start, seg_off;
    for (size_t i = 0; i < n; i++) {
        gm[i] = {ai.shift + in_off[i], in_len[i], out_off[i], out_cap[i], GRAPH_NO_SPLIT, 0};
        in_total = std::max<uint64_t>(in_total, ai.shift + in_off[i] + in_len[i]);
        mseg.push_back((uint32_t)seg_off.size());
        // segment boundaries on 16-byte addresses (full aligned frames
        // but for a member's first and last segment)
        const uint64_t lead = ((uintptr_t)d_out + out_off[i]) & 15;
        for (uint64_t o = 0; o < out_cap[i]; o += o ? CRC_SEG : CRC_SEG - lead) {
            seg_member.push_back((uint32_t)i);
            seg_start.push_back(o);
            seg_off.push_back(out_off[i] + o);
        }
    }
    g->in_total = in_total;
    mseg.push_back((uint32_t)seg_off.size());
    if (seg_off.size() > 0xFFFFFFFFull) throw StatusError{GZGPU_INVALID_ARGUMENT, "too many CRC segments"};
    g->nseg = (uint32_t)seg_off.size();
    auto put = [&](DevBuf &b, const void *src, size_t bytes) {
        b.ensure(std::max<size_t>(bytes, 16));
        if (bytes) HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    put(g->members, gm.data(), n * sizeof(GraphMember));

    // ---- which members split, and into how many pieces (decode cost from
    // the compressed length and the trailer ISIZE of the current contents)
    std::vector<uint64_t> cost(n, 1);
    std::vector<uint32_t> np(n, 1);
    if (split) {
        DevBuf isz;
        isz.ensure(std::max<size_t>(n * 4, 16));
        hipLaunchKernelGGL(k_graph_isize, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, c->stream, g->in_base,
                           g->members.as<GraphMember>(), (uint32_t)n, isz.as<uint32_t>());
        HIPCHK(hipGetLastError());
        std::vector<uint32_t> hisz(n);
        HIPCHK(hipMemcpyAsync(hisz.data(), isz.p, n * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        isz.release();
        const uint64_t slots =
            (uint64_t)c->num_cus *
            (c->tune.waves_per_cu ? std::min<uint32_t>(c->tune.waves_per_cu, inflate_waves_per_cu<uint16_t>(c, 9))
                                  : inflate_waves_per_cu<uint16_t>(c, 9));
        uint64_t total = 0;
        for (size_t i = 0; i < n; i++) {
            cost[i] = in_len[i] >= 18 ? member_cost((in_len[i] - 18) * 8, hisz[i]) : 1;
            total += cost[i];
        }
        const uint64_t target = std::max<uint64_t>(
            MIN_PIECE_OUT, (uint64_t)((double)total / (PIECES_PER_SLOT * SLOT_FRACTION * (double)slots)));
        for (size_t i = 0; i < n; i++) {
            if (in_len[i] < 1 * KIB || (c->tune.batch_split_bytes && in_len[i] <= c->tune.batch_split_bytes)) continue;
            // spacing floor as in the host planner: 2 KiB of compressed
            // input, down to 1 KiB for members whose output per bit is high
            const uint64_t cbits = (in_len[i] - 18) * 8;
            const uint64_t min_bits = std::min<uint64_t>(
                8 * MIN_SPACING_KIB * KIB,
                std::max<uint64_t>(8 * MIN_SPACING_FLOOR_KIB * KIB,
                                   (uint64_t)((double)MIN_PIECE_OUT * (double)cbits / (double)cost[i])));
            uint64_t k = (cost[i] + target - 1) / target;
            k = std::min<uint64_t>(k, cbits / min_bits);
            k = std::min<uint64_t>(k, 1u << 20);
            if (k >= 2) np[i] = (uint32_t)k;
        }
        g->miss_every = c->tune.miss_every;
    }
    std::vector<GraphSplit> sp;
    std::vector<uint32_t> pos_split, slot_pos;
    uint32_t next_piece = (uint32_t)n;
    uint64_t pool = 0, tiles = 0, nbits = 0;
    for (size_t i = 0; i < n; i++) {
        if (np[i] < 2) continue;
        GraphSplit s{};
        s.member = (uint32_t)i;
        s.np = np[i];
        s.first = next_piece;
        next_piece += np[i] - 1;
        s.slot0 = g->nslots;
        g->nslots += np[i];
        s.cell0 = g->ncells;
        g->ncells += np[i];
        s.pos0 = g->npos;
        g->npos += np[i] + 1;
        // marker slabs: estimates scaled to at most out_cap entries, x
        // GRAPH_SLAB_FACTOR + 4096 each, u16 entries (graph_slab_bytes),
        // rounded per piece
        s.slab_off = pool;
        s.slab_bytes = round_up((uint64_t)(2.0 * GRAPH_SLAB_FACTOR * (double)out_cap[i]) +
                                    (uint64_t)np[i] * (2 * (4096 + 65536 + 1) + 768) + 4096,
                                256);
        pool += s.slab_bytes;
        // re-decode slabs, each for the whole member: an overflowed piece's
        // second try (first half) and the remainder (second half)
        s.redo_off = pool;
        s.redo_bytes = marker_slab_bytes(out_cap[i]);
        pool += 2 * s.redo_bytes;
        tiles += out_cap[i] / RESOLVE_TILE + np[i] + 1;
        nbits += (in_len[i] - 18) * 8;
        g->max_chain = std::max(g->max_chain, np[i]);
        for (uint32_t k = 0; k <= np[i]; k++) pos_split.push_back((uint32_t)sp.size());
        for (uint32_t k = 0; k < np[i]; k++) slot_pos.push_back(k);
        gm[i].split = (uint32_t)sp.size();
        sp.push_back(s);
    }
    g->nsplit = (uint32_t)sp.size();
    g->npieces1 = next_piece;
    for (auto &s : sp) s.redo = next_piece++;
    if (g->nsplit) {
        if (in_total * 8 >= (1ull << 37)) throw StatusError{GZGPU_INVALID_ARGUMENT, "graph: input above 16 GiB"};
        put(g->members, gm.data(), n * sizeof(GraphMember));  // with the split indices
        put(g->splits, sp.data(), sp.size() * sizeof(GraphSplit));
        g->runs.ensure(sp.size() * sizeof(GraphSplitRun));
        g->cells.ensure((size_t)g->ncells * sizeof(GraphCell));
        g->acc.ensure((size_t)g->ncells * sizeof(GraphAcc));
        // finder ranges (bits of each split member after its shortest header,
        // before its trailer), sorted by start; hits map back to the member
        std::vector<uint32_t> ord(sp.size());
        for (uint32_t s = 0; s < ord.size(); s++) ord[s] = s;
        std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return gm[sp[a].member].in_off < gm[sp[b].member].in_off; });
        std::vector<BfRange> table;
        std::vector<uint64_t> lo, hi;
        std::vector<uint32_t> idx;
        for (uint32_t s : ord) {
            const GraphMember &m = gm[sp[s].member];
            const uint64_t a = (m.in_off + 10) * 8, b = (m.in_off + m.in_len - 8) * 8;
            // the last FINAL_SCAN_BITS also report BFINAL = 1 headers: the
            // final block then takes starts inside it like any other
            const uint64_t tail0 = b - std::min<uint64_t>(b - a, FINAL_SCAN_BITS);
            const uint64_t r[2][2] = {{a, tail0}, {tail0, b}};
            for (int k = 0; k < 2; k++) {
                if (r[k][1] <= r[k][0]) continue;
                table.push_back({r[k][0], r[k][1], g->nchunk, (uint32_t)k, 0u});
                g->nchunk += (r[k][1] - (r[k][0] & ~31ull) + 31) / 32;
                lo.push_back(r[k][0]);
                hi.push_back(r[k][1]);
                idx.push_back(s);
            }
        }
        g->nranges = (uint32_t)table.size();
        put(g->ranges, table.data(), table.size() * sizeof(BfRange));
        put(g->sp_lo, lo.data(), lo.size() * 8);
        put(g->sp_hi, hi.data(), hi.size() * 8);
        put(g->sp_idx, idx.data(), idx.size() * 4);
        g->surv_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(nbits / 64, 4096), 1ull << 28);
        g->hit_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(nbits / 1024, 4096), 1ull << 26);
        g->surv.ensure((size_t)g->surv_cap * 8);
        g->hits.ensure((size_t)g->hit_cap * sizeof(gzgpu_boundary));
        // the hits' set for the stored-block chain test: a power of two, at
        // least twice the hits it can hold (capped; a full set only leaves
        // stored hits unchained)
        g->ht_size = 1024;
        while (g->ht_size < 2ull * g->hit_cap && g->ht_size < (1u << 25)) g->ht_size <<= 1;
        g->htab.ensure((size_t)g->ht_size * 8);
        put(g->pos_split, pos_split.data(), pos_split.size() * 4);
        g->plen.ensure((size_t)g->npos * 8);
        g->ppos.ensure((size_t)g->npos * 8);
        g->rps.ensure((size_t)g->nslots * sizeof(ResolvePiece));
        std::vector<uint32_t> ident(g->nslots);
        for (uint32_t s = 0; s < g->nslots; s++) ident[s] = s;
        put(g->slot_rp, ident.data(), ident.size() * 4);
        put(g->slot_pos, slot_pos.data(), slot_pos.size() * 4);
        g->tile_first.ensure(((size_t)g->nslots + 1) * 4);
        if (tiles > 0xFFFFFFFFull) throw StatusError{GZGPU_INVALID_ARGUMENT, "graph: too many resolve tiles"};
        g->tile_cap = (uint32_t)tiles;
        g->tile_crc.ensure(std::max<size_t>(tiles, 1) * 4);
        g->scan.ensure((size_t)g->nslots * WIN * 2 * 2);
        g->pool.ensure(std::max<uint64_t>(pool, 256));
        // the second and third launches' piece lists, written on the device
        g->order2.ensure((size_t)g->npieces1 * 4);
        g->order3.ensure((size_t)g->nsplit * 4);
    }
    // first launch: every piece, longest (expected) first: whole members by
    // their cost, a split member's pieces by an equal share of it
    std::vector<uint64_t> pcost(g->npieces1);
    for (size_t i = 0; i < n; i++) pcost[i] = cost[i] / np[i];
    for (auto &s : sp)
        for (uint32_t k = 1; k < s.np; k++) pcost[s.first + k - 1] = cost[s.member] / s.np;
    std::vector<uint32_t> order(g->npieces1);
    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return pcost[a] > pcost[b]; });
    if (!split) {  // the plain form: by compressed length, as before
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return in_len[a] > in_len[b]; });
    }
    put(g->order, order.data(), order.size() * 4);
    put(g->seg_member, seg_member.data(), seg_member.size() * 4);
    put(g->seg_start, seg_start.data(), seg_start.size() * 8);
    put(g->seg_off, seg_off.data(), seg_off.size() * 8);
    put(g->mseg, mseg.data(), mseg.size() * 4);
    const size_t ptotal = (size_t)g->npieces1 + g->nsplit;
    g->pieces.ensure(ptotal * sizeof(PieceDesc));
    g->pout.ensure(ptotal * sizeof(PieceOut));
    g->counters.ensure(std::max<size_t>(64, (size_t)n * 4));
    g->head.ensure(n * sizeof(GraphHead));
    g->seg_len.ensure(std::max<size_t>(g->nseg, 1) * 4);
    g->seg_tail.ensure(std::max<size_t>(g->nseg, 1) * 8);
    g->seg_crc.ensure(std::max<size_t>(g->nseg, 1) * 4);
    g->mlen.ensure(n * 8);
    g->mstat.ensure(n * 4);
    g->mcrc.ensure(n * 4);
    g->status.ensure(n * 4);
    g->olen.ensure(n * 8);
    g->ocrc.ensure(n * 4);
    ensure_crc_tables(c);
    HIPCHK(hipStreamSynchronize(c->stream));
}

// The decode of the layout on the context's stream.  plain == false: the
// graph's body (split members planned on the device, one persistent inflate
// launch over every piece, one over the re-decodes); plain == true: one
// inflate launch per member, one wavefront each, queued back to back with no
// host synchronisation in between (the plain one-launch-per-member form the
// batched forms are measured against).
void graph_enqueue(gzgpu_batch_graph *g, bool plain) {
    gzgpu_ctx *c = g->c;
    const uint32_t nb = (g->n + 255) / 256, sb = (std::max(g->nseg, 1u) + 255) / 256;
    uint32_t *cnt = g->counters.as<uint32_t>();
    hipLaunchKernelGGL(k_graph_clear,
                       dim3(std::max<uint32_t>(1, std::min<uint32_t>((std::max(g->ncells, g->ht_size) + 255) / 256,
                                                                     c->num_cus * 4))),
                       dim3(256), 0, c->stream, g->cells.as<GraphCell>(), g->ncells, cnt, g->htab.as<uint64_t>(),
                       g->ht_size);
    hipLaunchKernelGGL(k_graph_prep, dim3(nb), dim3(256), 0, c->stream, g->in_base, g->d_out,
                       g->members.as<GraphMember>(), g->n, g->pieces.as<PieceDesc>(), g->head.as<GraphHead>(),
                       g->splits.as<GraphSplit>(), g->runs.as<GraphSplitRun>());
    const uint32_t slots16 = (uint32_t)c->num_cus * inflate_waves_per_cu<uint16_t>(c, 9);
    if (!plain && g->nsplit) {
        // block finder over the split members
        const uint32_t na = (uint32_t)std::min<uint64_t>((g->nchunk + 255) / 256, (uint64_t)c->num_cus * 16);
        hipLaunchKernelGGL(k_bf_stage_a, dim3(std::max<uint32_t>(na, 1)), dim3(256), 0, c->stream, g->in_base, g->in_total,
                           g->ranges.as<BfRange>(), g->nranges, g->nchunk, g->surv.as<uint64_t>(), g->surv_cap,
                           cnt + GC_SURV, g->hits.as<gzgpu_boundary>(), g->hit_cap, cnt + GC_HITS);
        const uint32_t nbb = std::min<uint32_t>(g->surv_cap / 256 + 1, (uint32_t)c->num_cus * 16);
        hipLaunchKernelGGL(k_bf_stage_b, dim3(nbb), dim3(256), 0, c->stream, g->in_base, g->in_total,
                           g->surv.as<uint64_t>(), cnt + GC_SURV, g->surv_cap, g->hits.as<gzgpu_boundary>(), g->hit_cap,
                           cnt + GC_HITS);
        hipLaunchKernelGGL(k_bf_estimate, dim3(std::min<uint32_t>(g->hit_cap, (uint32_t)c->num_cus * 64)), dim3(64), 0,
                           c->stream, g->in_base, g->in_total, g->hits.as<gzgpu_boundary>(), cnt + GC_HITS, g->hit_cap);
        hipLaunchKernelGGL(k_graph_hash_hits, dim3(std::min<uint32_t>((g->hit_cap + 255) / 256, (uint32_t)c->num_cus * 8)),
                           dim3(256), 0, c->stream, g->hits.as<gzgpu_boundary>(), cnt, g->hit_cap,
                           g->htab.as<uint64_t>(), g->ht_size - 1);
        hipLaunchKernelGGL(k_graph_bucket, dim3(std::min<uint32_t>((g->hit_cap + 255) / 256, (uint32_t)c->num_cus * 8)),
                           dim3(256), 0, c->stream, g->hits.as<gzgpu_boundary>(), cnt, g->hit_cap,
                           g->sp_lo.as<uint64_t>(), g->sp_hi.as<uint64_t>(), g->sp_idx.as<uint32_t>(), g->nranges,
                           g->splits.as<GraphSplit>(), g->runs.as<GraphSplitRun>(), g->cells.as<GraphCell>(),
                           (const uint64_t *)g->htab.p, g->ht_size - 1);
        hipLaunchKernelGGL(k_graph_plan, dim3(g->nsplit), dim3(GRAPH_PLAN_THREADS), 0, c->stream,
                           g->splits.as<GraphSplit>(), g->runs.as<GraphSplitRun>(), g->cells.as<GraphCell>(),
                           g->acc.as<GraphAcc>(), g->hits.as<gzgpu_boundary>(), g->members.as<GraphMember>(),
                           g->pieces.as<PieceDesc>(), g->pool.as<uint8_t>(), cnt, g->miss_every);
        hipLaunchKernelGGL(k_inflate<uint16_t>, dim3(std::min<uint32_t>(g->npieces1, slots16)), dim3(64), 0, c->stream,
                           g->in_base, g->in_total, g->pieces.as<PieceDesc>(), g->order.as<uint32_t>(), g->npieces1,
                           cnt + GC_QUEUE1, g->pout.as<PieceOut>(), (const uint32_t *)nullptr, (uint64_t *)nullptr, 0.f);
        hipLaunchKernelGGL(k_graph_link_local, dim3(g->nsplit), dim3(GRAPH_PLAN_THREADS), 0, c->stream,
                           g->splits.as<GraphSplit>(), g->runs.as<GraphSplitRun>(), g->acc.as<GraphAcc>(),
                           g->pout.as<PieceOut>(), g->pieces.as<PieceDesc>(), g->pool.as<uint8_t>(), cnt,
                           g->order2.as<uint32_t>());
        hipLaunchKernelGGL(k_inflate<uint16_t>, dim3(std::min<uint32_t>(g->npieces1, slots16)), dim3(64), 0, c->stream,
                           g->in_base, g->in_total, g->pieces.as<PieceDesc>(), g->order2.as<uint32_t>(), g->npieces1,
                           cnt + GC_QUEUE2, g->pout.as<PieceOut>(), (const uint32_t *)(cnt + GC_N2), (uint64_t *)nullptr, 0.f);
        hipLaunchKernelGGL(k_graph_link_final, dim3(g->nsplit), dim3(GRAPH_PLAN_THREADS), 0, c->stream,
                           g->splits.as<GraphSplit>(), g->runs.as<GraphSplitRun>(), g->acc.as<GraphAcc>(),
                           g->pout.as<PieceOut>(), g->pieces.as<PieceDesc>(), g->pool.as<uint8_t>(), cnt,
                           g->order3.as<uint32_t>());
        hipLaunchKernelGGL(k_inflate<uint16_t>, dim3(std::min<uint32_t>(g->nsplit, slots16)), dim3(64), 0, c->stream,
                           g->in_base, g->in_total, g->pieces.as<PieceDesc>(), g->order3.as<uint32_t>(), g->nsplit,
                           cnt + GC_QUEUE3, g->pout.as<PieceOut>(), (const uint32_t *)(cnt + GC_N3), (uint64_t *)nullptr, 0.f);
        hipLaunchKernelGGL(k_graph_resolve_prep, dim3(1), dim3(GRAPH_PREP_THREADS), 0, c->stream,
                           g->splits.as<GraphSplit>(), g->nsplit, g->runs.as<GraphSplitRun>(), g->acc.as<GraphAcc>(),
                           g->pout.as<PieceOut>(), g->pieces.as<PieceDesc>(), g->members.as<GraphMember>(), g->d_out,
                           g->pos_split.as<uint32_t>(), g->npos, g->plen.as<uint64_t>(), g->ppos.as<uint64_t>(),
                           g->rps.as<ResolvePiece>(), g->tile_first.as<uint32_t>(), g->nslots, g->mstat.as<int32_t>(),
                           g->mlen.as<uint64_t>());
        // the window chain of every split member as a parallel prefix, then
        // the bodies
        const dim3 sg(std::min<uint32_t>(g->nslots, (uint32_t)c->num_cus * 8));
        uint16_t *wb[2] = {g->scan.as<uint16_t>(), g->scan.as<uint16_t>() + (size_t)g->nslots * WIN};
        launch_chain_scan(c->stream, sg, (const ResolvePiece *)g->rps.p, g->slot_rp.as<uint32_t>(),
                          g->slot_pos.as<uint32_t>(), g->nslots, g->max_chain, wb);
        hipLaunchKernelGGL(k_resolve_body, dim3(std::min<uint32_t>(std::max<uint32_t>(g->tile_cap, 1), (uint32_t)c->num_cus * 32)),
                           dim3(256), 0, c->stream, (const ResolvePiece *)g->rps.p, g->tile_first.as<uint32_t>(),
                           g->nslots, c->d_crc_tab.as<uint32_t>(), g->tile_crc.as<uint32_t>(), 0u, 0xFFFFFFFFu);
    } else if (!plain) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(g->n, (uint64_t)c->num_cus * inflate_waves_per_cu<uint8_t>(c, 9));
        hipLaunchKernelGGL(k_inflate<uint8_t>, dim3(grid), dim3(64), 0, c->stream, g->in_base, g->in_total,
                           g->pieces.as<PieceDesc>(), g->order.as<uint32_t>(), g->n, cnt + GC_QUEUE1,
                           g->pout.as<PieceOut>(), (const uint32_t *)nullptr, (uint64_t *)nullptr, 0.f);
    } else {
        // member i: piece i, a queue counter of its own (cleared first), order
        // entry 0 of the identity (the kernel decodes pieces[order[0]])
        HIPCHK(hipMemsetAsync(g->counters.p, 0, (size_t)g->n * 4, c->stream));
        for (uint32_t i = 0; i < g->n; i++)
            hipLaunchKernelGGL(k_inflate<uint8_t>, dim3(1), dim3(64), 0, c->stream, g->in_base, g->in_total,
                               g->pieces.as<PieceDesc>() + i, g->zero.as<uint32_t>(), 1u, g->counters.as<uint32_t>() + i,
                               g->pout.as<PieceOut>() + i, (const uint32_t *)nullptr, (uint64_t *)nullptr, 0.f);
    }
    hipLaunchKernelGGL(k_graph_members, dim3(nb), dim3(256), 0, c->stream, g->members.as<GraphMember>(), g->n,
                       g->pout.as<PieceOut>(), g->mstat.as<int32_t>(), g->mlen.as<uint64_t>());
    hipLaunchKernelGGL(k_graph_seglen, dim3(sb), dim3(256), 0, c->stream, g->mlen.as<uint64_t>(),
                       g->seg_member.as<uint32_t>(), g->seg_start.as<uint64_t>(), g->nseg, g->seg_len.as<uint32_t>(),
                       g->seg_tail.as<uint64_t>());
    if (g->nseg)
        hipLaunchKernelGGL(k_crc_segments, dim3(std::min<uint32_t>(g->nseg, (uint32_t)c->num_cus * 8)), dim3(256), 0,
                           c->stream, (const uint8_t *)g->d_out, g->seg_off.as<uint64_t>(), g->seg_len.as<uint32_t>(),
                           g->seg_tail.as<uint64_t>(), g->nseg, c->d_crc_tab.as<uint32_t>(), g->seg_crc.as<uint32_t>());
    hipLaunchKernelGGL(k_crc_members, dim3(std::min<uint32_t>(g->n, (uint32_t)c->num_cus * 8)), dim3(256), 0,
                       c->stream, g->seg_crc.as<uint32_t>(), g->mseg.as<uint32_t>(), g->mlen.as<uint64_t>(), g->n,
                       g->mcrc.as<uint32_t>(), nullptr, nullptr);
    hipLaunchKernelGGL(k_graph_verify, dim3(nb), dim3(256), 0, c->stream, g->mstat.as<int32_t>(),
                       g->mlen.as<uint64_t>(), g->head.as<GraphHead>(), g->mcrc.as<uint32_t>(), g->n,
                       g->status.as<int32_t>(), g->olen.as<uint64_t>(), g->ocrc.as<uint32_t>());
}

// per-member results back to the caller; the first failing member's status
int graph_results(gzgpu_batch_graph *g, int32_t *status, uint64_t *out_len, uint32_t *crc, bool plain) {
    gzgpu_ctx *c = g->c;
    const int32_t *hs = download<int32_t>(c, g->status.p, g->n);
    const uint64_t *hl = download<uint64_t>(c, g->olen.p, g->n);
    const uint32_t *hc = download<uint32_t>(c, g->ocrc.p, g->n);
    const uint32_t *hk = plain ? nullptr : download<uint32_t>(c, g->counters.p, GC_N);
    HIPCHK(hipStreamSynchronize(c->stream));
    int first_bad = GZGPU_OK;
    uint64_t total = 0;
    for (uint32_t i = 0; i < g->n; i++) {
        if (status) status[i] = hs[i];
        if (out_len) out_len[i] = hl[i];
        if (crc) crc[i] = hc[i];
        total += hl[i];
        if (hs[i] != GZGPU_OK && first_bad == GZGPU_OK) {
            first_bad = hs[i];
            set_err(c, "member " + std::to_string(i) + ": " + gzgpu_status_string(hs[i]));
        }
    }
    if (!plain && g->nsplit && debug_level() >= 2) {
        // per split member: pieces, the first piece the link check refused and why
        std::vector<GraphSplit> hsp(g->nsplit);
        std::vector<GraphSplitRun> hrun(g->nsplit);
        HIPCHK(hipMemcpy(hsp.data(), g->splits.p, g->nsplit * sizeof(GraphSplit), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hrun.data(), g->runs.p, g->nsplit * sizeof(GraphSplitRun), hipMemcpyDeviceToHost));
        // the first launch's slowest pieces (wall clock at 100 MHz)
        std::vector<PieceOut> ho(g->npieces1);
        std::vector<PieceDesc> hd(g->npieces1);
        HIPCHK(hipMemcpy(ho.data(), g->pout.p, ho.size() * sizeof(PieceOut), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hd.data(), g->pieces.p, hd.size() * sizeof(PieceDesc), hipMemcpyDeviceToHost));
        std::vector<uint32_t> by(ho.size());
        for (uint32_t i = 0; i < by.size(); i++) by[i] = i;
        auto dur = [&](uint32_t i) { return ho[i].t_end > ho[i].t_begin ? ho[i].t_end - ho[i].t_begin : 0; };
        std::sort(by.begin(), by.end(), [&](uint32_t a, uint32_t b) { return dur(a) > dur(b); });
        for (uint32_t k = 0; k < std::min<uint32_t>(5, (uint32_t)by.size()); k++) {
            const uint32_t i = by[k];
            fprintf(stderr, "[gzgpu] graph slow piece %u (member %u, flags %u): %.3f ms, %llu entries, %llu bits, "
                            "%u blocks, status %u\n",
                    i, hd[i].member, hd[i].flags, (double)dur(i) / 1e5, (unsigned long long)(ho[i].n16 + ho[i].n8),
                    (unsigned long long)ho[i].bits_done, ho[i].nblocks, ho[i].status);
        }
        // pieces the first link check refused (re-decoded in place)
        std::vector<GraphAcc> ha(g->ncells);
        HIPCHK(hipMemcpy(ha.data(), g->acc.p, ha.size() * sizeof(GraphAcc), hipMemcpyDeviceToHost));
        for (uint32_t s = 0; s < g->nsplit; s++)
            for (uint32_t c2 = 1; c2 < hsp[s].np; c2++) {
                const GraphAcc &a = ha[hsp[s].cell0 + c2];
                if (!a.used || a.linked) continue;
                const uint32_t pi = hsp[s].first + c2 - 1;
                const uint32_t pp = a.pred ? hsp[s].first + a.pred - 1 : hsp[s].member;
                fprintf(stderr, "[gzgpu] graph miss: member %u cell %u/%u (pred cell %u): planned %llu (%s), start now %llu, "
                                "pred end %llu status %u; piece status %u err %u, %llu entries\n",
                        hsp[s].member, c2, hsp[s].np, a.pred, (unsigned long long)a.plan_lo,
                        a.plan_mid ? "inside a block" : "hit", (unsigned long long)a.lo,
                        (unsigned long long)ho[pp].end_bit, ho[pp].status, ho[pi].status, ho[pi].err,
                        (unsigned long long)(ho[pi].n16 + ho[pi].n8));
            }
        for (uint32_t s = 0; s < g->nsplit; s++)
            if (hrun[s].first_bad < hsp[s].np || debug_level() >= 3)
                fprintf(stderr, "[gzgpu] graph member %u: cells %u, first refused %u (%s), status %u, re-decoded %u\n",
                        hsp[s].member, hsp[s].np, hrun[s].first_bad,
                        hrun[s].why == 1 ? "unlinked" : hrun[s].why == 2 ? "overflow" : hrun[s].why == 3 ? "error" : "-",
                        hrun[s].fail, hrun[s].redo);
    }
    c->st.decompressed_bytes = total;
    // pieces: whole members plus the split members' marker pieces
    c->st.pieces = g->n + (hk ? hk[GC_PIECES] : 0);
    c->st.redecoded = hk ? hk[GC_REDO] : 0;
    c->st.candidates = hk && g->nsplit ? std::min<uint32_t>(hk[GC_HITS], g->hit_cap) : 0;
    return first_bad;
}

}  // namespace

extern "C" {

int gzgpu_batch_graph_create(gzgpu_ctx *c, const void *d_in, const uint64_t *in_off, const uint64_t *in_len,
                             void *d_out, const uint64_t *out_off, const uint64_t *out_cap, size_t n,
                             gzgpu_batch_graph **out) {
    if (!c || !out || !n || !d_in || !d_out || !in_off || !in_len || !out_off || !out_cap || n > 0xFFFFFFFFu)
        return GZGPU_INVALID_ARGUMENT;
    *out = nullptr;
    std::unique_ptr<gzgpu_batch_graph> g(new gzgpu_batch_graph());
    g->c = c;
    g->n = (uint32_t)n;
    const int st = guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        graph_setup(g.get(), d_in, in_off, in_len, d_out, out_off, out_cap, n, true);
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        graph_enqueue(g.get(), false);
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(c->stream, &g->graph);
        HIPCHK(le);
        HIPCHK(ce);
        HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        return (int)GZGPU_OK;
    });
    if (st != GZGPU_OK) {
        g->release();
        return st;
    }
    *out = g.release();
    return GZGPU_OK;
}

int gzgpu_decompress_members_plain(gzgpu_ctx *c, const void *d_in, const uint64_t *in_off, const uint64_t *in_len,
                                   void *d_out, const uint64_t *out_off, const uint64_t *out_cap, size_t n,
                                   int32_t *status, uint64_t *out_len, uint32_t *crc, gzgpu_stats *stats) {
    if (!c || !n || !d_in || !d_out || !in_off || !in_len || !out_off || !out_cap || n > 0xFFFFFFFFu)
        return GZGPU_INVALID_ARGUMENT;
    gzgpu_batch_graph g;
    g.c = c;
    g.n = (uint32_t)n;
    const int st = guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        c->st = gzgpu_stats{};
        graph_setup(&g, d_in, in_off, in_len, d_out, out_off, out_cap, n, false);
        const uint32_t zero = 0;
        g.zero.ensure(16);
        HIPCHK(hipMemcpy(g.zero.p, &zero, 4, hipMemcpyHostToDevice));
        HIPCHK(hipEventRecord(c->ev[0], c->stream));
        graph_enqueue(&g, true);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[4], c->stream));
        const int r = graph_results(&g, status, out_len, crc, true);
        c->st.ms_total = ev_ms(c->ev[0], c->ev[4]);
        c->st.launches = g.n + 7;
        c->st.inflate_launches = g.n;
        if (stats) *stats = c->st;
        return r;
    });
    g.release();
    return st;
}

int gzgpu_batch_graph_launch(gzgpu_batch_graph *g, int32_t *status, uint64_t *out_len, uint32_t *crc,
                             gzgpu_stats *stats) {
    if (!g || !g->c) return GZGPU_INVALID_ARGUMENT;
    gzgpu_ctx *c = g->c;
    return guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        c->st = gzgpu_stats{};
        HIPCHK(hipEventRecord(c->ev[0], c->stream));
        HIPCHK(hipGraphLaunch(g->exec, c->stream));
        HIPCHK(hipEventRecord(c->ev[4], c->stream));
        const int first_bad = graph_results(g, status, out_len, crc, false);
        c->st.ms_total = ev_ms(c->ev[0], c->ev[4]);
        c->st.launches = 1;
        c->st.inflate_launches = g->nsplit ? 2 : 1;
        if (stats) *stats = c->st;
        return first_bad;
    });
}

void gzgpu_batch_graph_destroy(gzgpu_batch_graph *g) {
    if (!g) return;
    if (g->c) {
        (void)hipSetDevice(g->c->device);
        if (g->c->stream) (void)hipStreamSynchronize(g->c->stream);
    }
    g->release();
    delete g;
}

}  // extern "C"
