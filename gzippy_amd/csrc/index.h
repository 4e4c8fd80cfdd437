// index.h — seekable index: random access into a single-member gzip stream
// from checkpoints (src/decompress/index.rs: build_index :54,
// serialize_index :145, load_index :184, seek_decompress :344).
//
// Included at the end of gzgpu.hip (it drives the member pipeline).
//
// Build: one full decode of the member through the pipeline, which reports
// every block start a piece decodes (bit after the deflate start, output
// offset; a piece that starts inside a block reports the next block start).
// Checkpoints are the first of those, then every next one at least
// `interval` output bytes further; each carries the 32 KiB of output before
// it.  Serialized in index.rs's version-1 layout: "GZIDX\x01", deflate offset
// u32, total output u64, point count u32, 2 reserved bytes, then per point
// the bit offset u64 (from the deflate start), the output offset u64 and the
// 32768-byte window.
//
// Seek: the checkpoints around [offset, offset + max_bytes) decode as one
// batch of pieces with known windows (k_inflate<uint8_t>, one wavefront per
// stretch between consecutive checkpoints, all at once); each piece must
// stop exactly at the next checkpoint's bit with exactly the output the
// index says, or the index does not belong to the input.  Only the input
// bytes those pieces read go to the device.

namespace {

constexpr size_t INDEX_HEAD = 6 + 4 + 8 + 4 + 2;
constexpr size_t INDEX_POINT = 8 + 8 + WIN;
const uint8_t INDEX_MAGIC_V1[6] = {'G', 'Z', 'I', 'D', 'X', 1};

void put_le(uint8_t *p, uint64_t v, int n) {
    for (int i = 0; i < n; i++) p[i] = (uint8_t)(v >> (8 * i));
}
uint64_t get_le(const uint8_t *p, int n) {
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

}  // namespace

extern "C" {

int gzgpu_index_build(gzgpu_ctx *c, const uint8_t *in, size_t in_len, uint64_t interval, uint8_t *index, size_t cap,
                      size_t *index_len) {
    if (!c || (!in && in_len) || !index_len) return GZGPU_INVALID_ARGUMENT;
    *index_len = 0;
    return guarded(c, [&]() -> int {
        HIPCHK(hipSetDevice(c->device));
        c->st = gzgpu_stats{};
        if (in_len < 2 || in[0] != 0x1f || in[1] != 0x8b) return finish(c, GZGPU_INVALID_ARGUMENT, "not a gzip file");
        if (sized_member_walk(in, in_len).size() > 2)
            return finish(c, GZGPU_INVALID_ARGUMENT,
                          "members that carry their size (BGZF / gzippy parallel) have their own block boundaries");
        Member M;
        std::string msg;
        int st = member_from_host(in, in_len, 0, M, msg);
        if (st) return finish(c, st, msg);
        const uint64_t hl = M.ds / 8;
        c->d_in.ensure(in_len + 64);
        HIPCHK(hipMemcpyAsync(c->d_in.p, in, in_len, hipMemcpyHostToDevice, c->stream));
        c->d_out.ensure((uint64_t)M.exp_isize + 64);
        M.out = c->d_out.as<uint8_t>();
        M.out_cap = M.exp_isize;
        M.split = true;
        std::vector<std::pair<uint64_t, uint64_t>> cps;
        c->cp_collect = &cps;
        std::vector<Member> v{M};
        try {
            run_members(c, c->d_in.as<uint8_t>(), in_len, v);
        } catch (...) {
            c->cp_collect = nullptr;
            throw;
        }
        c->cp_collect = nullptr;
        M = v[0];
        if (M.status != GZGPU_OK) return finish(c, M.status, M.msg);
        if ((M.end_bit + 7) / 8 + 8 != in_len)
            return finish(c, GZGPU_INVALID_ARGUMENT, "multi-member gzip files are not supported for indexing");
        // checkpoints: the deflate start, then every block start at least
        // `interval` output bytes after the previous checkpoint
        std::sort(cps.begin(), cps.end(), [](const std::pair<uint64_t, uint64_t> &a, const std::pair<uint64_t, uint64_t> &b) {
            return a.second < b.second || (a.second == b.second && a.first < b.first);
        });
        std::vector<std::pair<uint64_t, uint64_t>> pts;
        pts.push_back({0, 0});
        for (auto &p : cps)
            if (p.second < M.out_len && p.first > pts.back().first && p.second >= pts.back().second + std::max<uint64_t>(interval, 1))
                pts.push_back(p);
        const size_t need = INDEX_HEAD + pts.size() * INDEX_POINT;
        *index_len = need;
        if (!index || cap < need) {
            set_err(c, "index buffer too small: " + std::to_string(need) + " bytes needed");
            return (int)GZGPU_OUTPUT_TOO_SMALL;
        }
        // windows: the 32 KiB of output before each point (zeros before the
        // stream start), straight from the decoded output
        uint8_t *w = (uint8_t *)c->pin.alloc(pts.size() * WIN);
        for (size_t i = 0; i < pts.size(); i++) {
            const uint64_t o = pts[i].second, have = std::min<uint64_t>(o, WIN);
            memset(w + i * WIN, 0, WIN - have);
            if (have)
                HIPCHK(hipMemcpyAsync(w + i * WIN + (WIN - have), c->d_out.as<uint8_t>() + o - have, have,
                                      hipMemcpyDeviceToHost, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        memcpy(index, INDEX_MAGIC_V1, 6);
        put_le(index + 6, hl, 4);
        put_le(index + 10, M.out_len, 8);
        put_le(index + 18, pts.size(), 4);
        put_le(index + 22, 0, 2);
        for (size_t i = 0; i < pts.size(); i++) {
            uint8_t *q = index + INDEX_HEAD + i * INDEX_POINT;
            put_le(q, pts[i].first, 8);
            put_le(q + 8, pts[i].second, 8);
            memcpy(q + 16, w + i * WIN, WIN);
        }
        return (int)GZGPU_OK;
    });
}

int gzgpu_index_seek(gzgpu_ctx *c, const uint8_t *in, size_t in_len, const uint8_t *index, size_t index_len,
                     uint64_t offset, uint64_t max_bytes, uint8_t *out, size_t out_cap, size_t *out_len) {
    if (!c || (!in && in_len) || !index || !out_len) return GZGPU_INVALID_ARGUMENT;
    *out_len = 0;
    return guarded(c, [&]() -> int {
        HIPCHK(hipSetDevice(c->device));
        c->st = gzgpu_stats{};
        if (in_len < 2 || in[0] != 0x1f || in[1] != 0x8b) return finish(c, GZGPU_INVALID_ARGUMENT, "not a gzip file");
        if (index_len < INDEX_HEAD || memcmp(index, INDEX_MAGIC_V1, 5) != 0)
            return finish(c, GZGPU_INVALID_ARGUMENT, "invalid index file magic");
        if (index[5] != 1)
            return finish(c, GZGPU_INVALID_ARGUMENT,
                          "index version " + std::to_string(index[5]) + ": only version 1 (bit offsets) is supported");
        const uint64_t hl = get_le(index + 6, 4), total = get_le(index + 10, 8), npts = get_le(index + 18, 4);
        if (index_len < INDEX_HEAD + npts * INDEX_POINT || npts == 0)
            return finish(c, GZGPU_INVALID_ARGUMENT, "index file truncated");
        auto pt_bit = [&](size_t i) { return get_le(index + INDEX_HEAD + i * INDEX_POINT, 8); };
        auto pt_out = [&](size_t i) { return get_le(index + INDEX_HEAD + i * INDEX_POINT + 8, 8); };
        auto pt_win = [&](size_t i) { return index + INDEX_HEAD + i * INDEX_POINT + 16; };
        for (size_t i = 1; i < npts; i++)
            if (pt_out(i) <= pt_out(i - 1) || pt_bit(i) <= pt_bit(i - 1))
                return finish(c, GZGPU_INVALID_ARGUMENT, "index points are not monotonically increasing");
        if (hl + 8 > in_len) return finish(c, GZGPU_INVALID_ARGUMENT, "index does not match the input");
        if (offset > total)
            return finish(c, GZGPU_INVALID_ARGUMENT, "offset " + std::to_string(offset) + " is beyond uncompressed size " +
                                                         std::to_string(total));
        const uint64_t end = std::min<uint64_t>(total, offset + std::min<uint64_t>(max_bytes, total));
        if (end == offset) return (int)GZGPU_OK;
        // checkpoints [i, j): i holds offset, j the first at or past end
        size_t i = 0;
        while (i + 1 < npts && pt_out(i + 1) <= offset) i++;
        size_t j = i + 1;
        while (j < npts && pt_out(j) < end) j++;
        const uint64_t ds = hl * 8, de = (in_len - 8) * 8;
        // input bytes the pieces read: from the first point's byte to the
        // next point's byte (+ a margin for the stop header), or the end
        const uint64_t b0 = (ds + pt_bit(i)) / 8;
        const uint64_t b1 = j < npts ? std::min<uint64_t>(in_len, (ds + pt_bit(j)) / 8 + 64) : in_len;
        const uint64_t base = b0 & ~(uint64_t)15;
        c->d_in.ensure(b1 - base + 64);
        HIPCHK(hipMemcpyAsync(c->d_in.p, in + base, b1 - base, hipMemcpyHostToDevice, c->stream));
        const uint64_t span_out = (j < npts ? pt_out(j) : total) - pt_out(i);
        c->d_out.ensure(span_out + 64);
        c->d_env.ensure((j - i) * WIN);
        uint8_t *hw = (uint8_t *)c->pin.alloc((j - i) * WIN);
        for (size_t q = i; q < j; q++) memcpy(hw + (q - i) * WIN, pt_win(q), WIN);
        HIPCHK(hipMemcpyAsync(c->d_env.p, hw, (j - i) * WIN, hipMemcpyHostToDevice, c->stream));
        std::vector<PieceDesc> d;
        std::vector<uint64_t> cost;
        const uint64_t shift = base * 8;
        for (size_t q = i; q < j; q++) {
            PieceDesc pd{};
            pd.start_bit = ds + pt_bit(q) - shift;
            pd.stop_bit = q + 1 < npts ? ds + pt_bit(q + 1) - shift : NO_STOP;
            pd.limit_bit = (q + 1 < npts ? std::min<uint64_t>(de, (b1 - 0) * 8) : de) - shift;
            pd.dst = c->d_out.as<uint8_t>() + (pt_out(q) - pt_out(i));
            pd.cap = (q + 1 < npts ? pt_out(q + 1) : total) - pt_out(q);
            pd.window = c->d_env.as<uint8_t>() + (q - i) * WIN;
            pd.flags = PF_WINDOW;
            pd.member = 0;
            d.push_back(pd);
            cost.push_back(pd.cap + 1);
        }
        float ms = 0;
        const std::vector<PieceOut> r =
            finish_inflate(c, enqueue_inflate<uint8_t>(c, c->d_in.as<uint8_t>(), b1 - base, d, cost), d, ms);
        for (size_t q = i; q < j; q++) {
            const PieceOut &o = r[q - i];
            const uint64_t want = (q + 1 < npts ? pt_out(q + 1) : total) - pt_out(q);
            const bool at_stop = q + 1 < npts ? (o.status == PS_STOPPED && o.end_bit == d[q - i].stop_bit)
                                              : o.status == PS_FINAL;
            if (!at_stop || o.n8 != want)
                return finish(c, GZGPU_DECODE_FAILED,
                              "checkpoint " + std::to_string(q) + " does not decode as the index says (the index "
                              "belongs to another input)");
        }
        const uint64_t n = end - offset;
        if (n > out_cap) {
            *out_len = n;
            set_err(c, "output buffer too small");
            return (int)GZGPU_OUTPUT_TOO_SMALL;
        }
        HIPCHK(hipMemcpyAsync(out, c->d_out.as<uint8_t>() + (offset - pt_out(i)), n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        *out_len = n;
        return (int)GZGPU_OK;
    });
}

}  // extern "C"
