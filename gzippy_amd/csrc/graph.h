// graph.h — a batch of members as one replayable hipGraph.
//
// Included at the end of gzgpu.hip (same translation unit).
//
// gzgpu_batch_graph_create fixes a layout: where each member's compressed
// bytes lie in d_in and where its output goes in d_out.  It captures, on the
// context's stream, every step of decoding that batch with no host decision
// in between:
//   1. k_graph_prep: one thread per member parses the gzip header from device
//      memory (gzip_format.rs:47), reads the trailer, and writes a DIRECT
//      piece descriptor (a bad header gets a descriptor that fails at once);
//      it also clears the inflate work queue (a kernel, not a memset node);
//   2. k_inflate<uint8_t>: the persistent kernel pulls members from the
//      device-side queue, longest first (the order is fixed at creation from
//      the compressed lengths), one wavefront each;
//   3. k_graph_seglen: CRC segment lengths from each member's decoded length;
//   4. k_crc_segments + k_crc_members: CRC32 per member;
//   5. k_graph_verify: status per member (decode, ISIZE, CRC32 against the
//      trailer).
// gzgpu_batch_graph_launch replays it with one hipGraphLaunch and copies the
// per-member status, length and CRC back.  The caller may rewrite the
// contents of d_in between launches; offsets, lengths and capacities stay.
// Members are decoded whole, one wavefront each: the form suits batches of
// many small members (gzgpu_decompress_batch splits large ones).

namespace gz {

struct GraphMember {
    uint64_t in_off, in_len;    // bytes, relative to the aligned input base
    uint64_t out_off, out_cap;  // bytes, relative to d_out
};

struct GraphHead {
    uint32_t status;  // GZGPU_OK or the header error
    uint32_t exp_crc, exp_isize, pad;
};

__global__ void k_graph_prep(const uint8_t *__restrict__ in, uint8_t *out, const GraphMember *__restrict__ gm,
                             uint32_t n, PieceDesc *__restrict__ pieces, GraphHead *__restrict__ head,
                             uint32_t *queue) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) queue[0] = 0;
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
}

// Segment s of member seg_member[s] covers its bytes from seg_start[s] to
// the next segment's start (64 KiB for its last), clipped to what the member
// decoded; seg_tail is what the member decoded after it.
__global__ void k_graph_seglen(const PieceOut *__restrict__ pout, const uint32_t *__restrict__ seg_member,
                               const uint64_t *__restrict__ seg_start, uint32_t nseg, uint32_t n,
                               uint32_t *__restrict__ seg_len, uint64_t *__restrict__ seg_tail,
                               uint64_t *__restrict__ mlen) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nseg) {
        const PieceOut &r = pout[seg_member[i]];
        const uint64_t got = r.status == PS_FINAL ? r.n8 : 0;
        const uint64_t s = seg_start[i];
        const uint64_t cap = i + 1 < nseg && seg_member[i + 1] == seg_member[i] ? seg_start[i + 1] - s : CRC_SEG;
        const uint32_t l = got > s ? (uint32_t)(got - s < cap ? got - s : cap) : 0u;
        seg_len[i] = l;
        seg_tail[i] = got > s + l ? got - s - l : 0;
    }
    if (i < n) mlen[i] = pout[i].status == PS_FINAL ? pout[i].n8 : 0;
}

__global__ void k_graph_verify(const PieceOut *__restrict__ pout, const GraphHead *__restrict__ head,
                               const uint32_t *__restrict__ crc, uint32_t n, int32_t *__restrict__ status,
                               uint64_t *__restrict__ out_len, uint32_t *__restrict__ out_crc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GraphHead h = head[i];
    const PieceOut &r = pout[i];
    int32_t st = (int32_t)h.status;
    uint64_t len = 0;
    if (st == GZGPU_OK) {
        if (r.status == PS_FINAL) {
            len = r.n8;
            if ((uint32_t)len != h.exp_isize) st = GZGPU_SIZE_MISMATCH;
            else if (crc[i] != h.exp_crc) st = GZGPU_CRC_MISMATCH;
        } else {
            st = r.status == PS_OVERFLOW ? GZGPU_OUTPUT_TOO_SMALL : GZGPU_DECODE_FAILED;
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
    const uint8_t *in_base = nullptr;
    uint8_t *d_out = nullptr;
    uint64_t in_total = 0;
    DevBuf members, pieces, order, pout, queue, head, seg_member, seg_start, seg_off, seg_len, seg_tail, mseg, mlen,
        seg_crc, mcrc, status, olen, ocrc, zero;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        for (DevBuf *b : {&members, &pieces, &order, &pout, &queue, &head, &seg_member, &seg_start, &seg_off, &seg_len,
                          &seg_tail, &mseg, &mlen, &seg_crc, &mcrc, &status, &olen, &ocrc, &zero})
            b->release();
    }
};

namespace {

// Buffers of a fixed batch layout (members, their CRC segments, longest-first
// order), uploaded once.
void graph_setup(gzgpu_batch_graph *g, const void *d_in, const uint64_t *in_off, const uint64_t *in_len,
                 void *d_out, const uint64_t *out_off, const uint64_t *out_cap, size_t n) {
    gzgpu_ctx *c = g->c;
    const AlignedIn ai = align_input(d_in);
    g->in_base = ai.base;
    g->d_out = (uint8_t *)d_out;
    uint64_t in_total = 0;
    std::vector<GraphMember> gm(n);
    std::vector<uint32_t> seg_member, mseg;
    std::vector<uint64_t> seg_start, seg_off;
    for (size_t i = 0; i < n; i++) {
        gm[i] = {ai.shift + in_off[i], in_len[i], out_off[i], out_cap[i]};
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
    // longest compressed member first
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return in_len[a] > in_len[b]; });
    auto put = [&](DevBuf &b, const void *src, size_t bytes) {
        b.ensure(std::max<size_t>(bytes, 16));
        if (bytes) HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    put(g->members, gm.data(), n * sizeof(GraphMember));
    put(g->order, order.data(), n * 4);
    put(g->seg_member, seg_member.data(), seg_member.size() * 4);
    put(g->seg_start, seg_start.data(), seg_start.size() * 8);
    put(g->seg_off, seg_off.data(), seg_off.size() * 8);
    put(g->mseg, mseg.data(), mseg.size() * 4);
    g->pieces.ensure(n * sizeof(PieceDesc));
    g->pout.ensure(n * sizeof(PieceOut));
    g->queue.ensure(std::max<size_t>(64, n * 4));
    g->head.ensure(n * sizeof(GraphHead));
    g->seg_len.ensure(std::max<size_t>(g->nseg, 1) * 4);
    g->seg_tail.ensure(std::max<size_t>(g->nseg, 1) * 8);
    g->seg_crc.ensure(std::max<size_t>(g->nseg, 1) * 4);
    g->mlen.ensure(n * 8);
    g->mcrc.ensure(n * 4);
    g->status.ensure(n * 4);
    g->olen.ensure(n * 8);
    g->ocrc.ensure(n * 4);
    ensure_crc_tables(c);
    HIPCHK(hipStreamSynchronize(c->stream));
}

// The decode of the layout on the context's stream.  plain == false: one
// persistent inflate launch over every member (the graph's body);
// plain == true: one inflate launch per member, one wavefront each, queued
// back to back with no host synchronisation in between (the plain
// one-launch-per-member form the batched forms are measured against).
void graph_enqueue(gzgpu_batch_graph *g, bool plain) {
    gzgpu_ctx *c = g->c;
    const uint32_t nb = (g->n + 255) / 256, sb = (std::max(g->nseg, g->n) + 255) / 256;
    hipLaunchKernelGGL(k_graph_prep, dim3(nb), dim3(256), 0, c->stream, g->in_base, g->d_out,
                       g->members.as<GraphMember>(), g->n, g->pieces.as<PieceDesc>(), g->head.as<GraphHead>(),
                       g->queue.as<uint32_t>());
    if (!plain) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(g->n, (uint64_t)c->num_cus * inflate_waves_per_cu<uint8_t>(c, 9));
        hipLaunchKernelGGL(k_inflate<uint8_t>, dim3(grid), dim3(64), 0, c->stream, g->in_base, g->in_total,
                           g->pieces.as<PieceDesc>(), g->order.as<uint32_t>(), g->n, g->queue.as<uint32_t>(),
                           g->pout.as<PieceOut>());
    } else {
        // member i: piece i, a queue counter of its own (cleared first), order
        // entry 0 of the identity (the kernel decodes pieces[order[0]])
        HIPCHK(hipMemsetAsync(g->queue.p, 0, (size_t)g->n * 4, c->stream));
        for (uint32_t i = 0; i < g->n; i++)
            hipLaunchKernelGGL(k_inflate<uint8_t>, dim3(1), dim3(64), 0, c->stream, g->in_base, g->in_total,
                               g->pieces.as<PieceDesc>() + i, g->zero.as<uint32_t>(), 1u, g->queue.as<uint32_t>() + i,
                               g->pout.as<PieceOut>() + i);
    }
    hipLaunchKernelGGL(k_graph_seglen, dim3(sb), dim3(256), 0, c->stream, g->pout.as<PieceOut>(),
                       g->seg_member.as<uint32_t>(), g->seg_start.as<uint64_t>(), g->nseg, g->n,
                       g->seg_len.as<uint32_t>(), g->seg_tail.as<uint64_t>(), g->mlen.as<uint64_t>());
    if (g->nseg)
        hipLaunchKernelGGL(k_crc_segments, dim3(std::min<uint32_t>(g->nseg, (uint32_t)c->num_cus * 8)), dim3(256), 0,
                           c->stream, (const uint8_t *)g->d_out, g->seg_off.as<uint64_t>(), g->seg_len.as<uint32_t>(),
                           g->seg_tail.as<uint64_t>(), g->nseg, c->d_crc_tab.as<uint32_t>(), g->seg_crc.as<uint32_t>());
    hipLaunchKernelGGL(k_crc_members, dim3(std::min<uint32_t>(g->n, (uint32_t)c->num_cus * 8)), dim3(256), 0,
                       c->stream, g->seg_crc.as<uint32_t>(), g->mseg.as<uint32_t>(), g->mlen.as<uint64_t>(), g->n,
                       g->mcrc.as<uint32_t>());
    hipLaunchKernelGGL(k_graph_verify, dim3(nb), dim3(256), 0, c->stream, g->pout.as<PieceOut>(),
                       g->head.as<GraphHead>(), g->mcrc.as<uint32_t>(), g->n, g->status.as<int32_t>(),
                       g->olen.as<uint64_t>(), g->ocrc.as<uint32_t>());
}

// per-member results back to the caller; the first failing member's status
int graph_results(gzgpu_batch_graph *g, int32_t *status, uint64_t *out_len, uint32_t *crc) {
    gzgpu_ctx *c = g->c;
    const int32_t *hs = download<int32_t>(c, g->status.p, g->n);
    const uint64_t *hl = download<uint64_t>(c, g->olen.p, g->n);
    const uint32_t *hc = download<uint32_t>(c, g->ocrc.p, g->n);
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
    c->st.decompressed_bytes = total;
    c->st.pieces = g->n;
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
        graph_setup(g.get(), d_in, in_off, in_len, d_out, out_off, out_cap, n);
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
        graph_setup(&g, d_in, in_off, in_len, d_out, out_off, out_cap, n);
        const uint32_t zero = 0;
        g.zero.ensure(16);
        HIPCHK(hipMemcpy(g.zero.p, &zero, 4, hipMemcpyHostToDevice));
        HIPCHK(hipEventRecord(c->ev[0], c->stream));
        graph_enqueue(&g, true);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[4], c->stream));
        const int r = graph_results(&g, status, out_len, crc);
        c->st.ms_total = ev_ms(c->ev[0], c->ev[4]);
        c->st.launches = g.n + 5;
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
        const int first_bad = graph_results(g, status, out_len, crc);
        c->st.ms_total = ev_ms(c->ev[0], c->ev[4]);
        c->st.launches = 1;
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
