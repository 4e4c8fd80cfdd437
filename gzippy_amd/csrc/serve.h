// serve.h — continuous-batching request server over the batch pipeline.
//
// Included at the end of gzgpu.hip (same translation unit: it drives the
// member pipeline directly).
//
// A request is one gzip stream, one member or several concatenated.
// gzgpu_submit copies it into host memory and queues it.  Each worker thread
// owns a context (its own HIP stream and device buffers) and loops:
//   1. admit: take every pending request, oldest first, while the batch stays
//      under max_batch_requests and the server under its HBM budget (input +
//      expected output, from the gzip trailers);
//   2. split each request into members at gzip magics (the reference's
//      member-boundary scan, scan_member_boundaries_fast bgzf.rs:2090, checked
//      by decoding: a span that is not exactly one member fails its own
//      trailer check);
//   3. one H2D copy of all inputs, one pass of the member pipeline over every
//      member of every request (one persistent inflate launch for all), one
//      D2H copy per request;
//   4. a request whose split was wrong is decoded again member by member
//      (gzippy::decompress semantics, decompress/mod.rs:725-741).
// With two workers, a batch uploads and decodes while the previous one is
// still on the device, and requests that arrive meanwhile form the next batch.

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct ServeRequest {
    uint64_t id = 0;
    std::vector<uint8_t> in;
    std::vector<uint8_t> out;
    int status = -1;  // -1: not finished
    uint32_t members = 0;
    std::string msg;
    uint64_t t_submit = 0, t_start = 0, t_done = 0;
    uint64_t batch = 0;
    uint64_t reserve = 0;  // HBM bytes reserved while admitted
};

}  // namespace

struct gzgpu_server {
    int device = 0;
    gzgpu_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<ServeRequest>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<ServeRequest>> live;
    uint64_t next_id = 1, next_batch = 1;
    uint64_t hbm_in_use = 0;
    bool stop = false;
    gzgpu_server_stats st{};
    std::vector<std::thread> workers;
    std::vector<gzgpu_ctx *> ctxs;
};

namespace {

uint64_t request_reserve(const ServeRequest &r) {
    // input (16-byte padded), expected output from the members' trailers, and
    // the speculative pieces' marker slabs (u16 entries until the flip, sized
    // 1.3 x the output share: up to 2.6 x the output)
    uint64_t out = 0;
    for (auto &sp : member_spans(r.in.data(), r.in.size())) {
        const uint64_t isz = span_isize(r.in.data() + sp.first, sp.second);
        out += isz == ~0ull ? 1032ull * sp.second : isz;
    }
    return (r.in.size() + 16) + out + (out * 26) / 10 + 64 * 1024;
}

// Decode one admitted batch on context c.
void serve_batch(gzgpu_server *S, gzgpu_ctx *c, std::vector<std::shared_ptr<ServeRequest>> &batch) {
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->pin.reset();
    c->st = gzgpu_stats{};
    // layout: request inputs back to back (16-byte aligned), members' outputs
    // back to back per request
    struct Span {
        uint32_t req;
        size_t off, len;  // within the request
    };
    std::vector<Span> spans;
    std::vector<uint64_t> in_base(batch.size()), out_base(batch.size()), out_size(batch.size());
    std::vector<uint8_t> req_bad(batch.size(), 0);
    uint64_t in_total = 0, out_total = 0;
    for (uint32_t i = 0; i < batch.size(); i++) {
        ServeRequest &r = *batch[i];
        in_base[i] = in_total;
        in_total += (r.in.size() + 15) & ~(uint64_t)15;
        out_base[i] = out_total;
        uint64_t osz = 0;
        auto v = member_spans(r.in.data(), r.in.size());
        r.members = (uint32_t)v.size();
        bool plausible = true;
        for (auto &sp : v) plausible = plausible && span_isize(r.in.data() + sp.first, sp.second) != ~0ull;
        if (!plausible && r.in.size() >= 2 && r.in[0] == 0x1f && r.in[1] == 0x8b) {
            // a false member boundary: straight to the member-by-member walk
            req_bad[i] = 1;
            out_size[i] = 0;
            continue;
        }
        for (auto &sp : v) {
            spans.push_back({i, sp.first, sp.second});
            if (sp.second >= 8) osz += le32(r.in.data() + sp.first + sp.second - 4);
        }
        out_size[i] = osz;
        out_total += (osz + 15) & ~(uint64_t)15;
    }
    c->d_in.ensure(std::max<uint64_t>(in_total, 16) + 64);
    c->d_out.ensure(std::max<uint64_t>(out_total, 16) + 64);
    uint8_t *hin = (uint8_t *)c->pin.alloc(std::max<uint64_t>(in_total, 16));
    for (uint32_t i = 0; i < batch.size(); i++)
        if (!batch[i]->in.empty()) memcpy(hin + in_base[i], batch[i]->in.data(), batch[i]->in.size());
    HIPCHK(hipMemcpyAsync(c->d_in.p, hin, in_total, hipMemcpyHostToDevice, c->stream));
    // one Member per span
    std::vector<Member> mem;
    std::vector<uint32_t> mem_req;
    {
        std::vector<uint64_t> out_next(out_base);
        for (auto &sp : spans) {
            ServeRequest &r = *batch[sp.req];
            Member M;
            std::string msg;
            if (r.in.size() < 2 || r.in[0] != 0x1f || r.in[1] != 0x8b) continue;  // non-gzip: empty output
            int st = member_from_host(r.in.data() + sp.off, sp.len, in_base[sp.req] + sp.off, M, msg);
            if (st) {
                req_bad[sp.req] = 1;
                continue;
            }
            M.out = c->d_out.as<uint8_t>() + out_next[sp.req];
            M.out_cap = M.exp_isize;
            out_next[sp.req] += M.exp_isize;
            M.split = true;
            mem.push_back(M);
            mem_req.push_back(sp.req);
        }
    }
    if (!mem.empty()) run_members(c, c->d_in.as<uint8_t>(), in_total, mem);
    for (size_t k = 0; k < mem.size(); k++) {
        const Member &M = mem[k];
        // the member must end exactly where its span's trailer starts
        const bool exact = M.status == GZGPU_OK && (M.end_bit + 7) / 8 == M.de / 8;
        if (!exact) req_bad[mem_req[k]] = 1;
    }
    // results: one D2H per good request
    std::vector<uint8_t *> hout(batch.size(), nullptr);
    for (uint32_t i = 0; i < batch.size(); i++) {
        if (req_bad[i] || out_size[i] == 0) continue;
        hout[i] = (uint8_t *)c->pin.alloc(out_size[i]);
        HIPCHK(hipMemcpyAsync(hout[i], c->d_out.as<uint8_t>() + out_base[i], out_size[i], hipMemcpyDeviceToHost,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    uint64_t fallbacks = 0;
    for (uint32_t i = 0; i < batch.size(); i++) {
        ServeRequest &r = *batch[i];
        r.out.clear();
        if (r.in.size() < 2 || r.in[0] != 0x1f || r.in[1] != 0x8b) {
            r.status = GZGPU_OK;  // not gzip: empty result, like decompress()
            continue;
        }
        if (!req_bad[i]) {
            r.out.assign(hout[i], hout[i] + out_size[i]);
            r.status = GZGPU_OK;
            continue;
        }
        // split miss or a bad member: the request alone, member by member
        fallbacks++;
        uint64_t total = 0;
        int st = decompress_host_stream(
            c, r.in.data(), r.in.size(),
            [&](const uint8_t *d, size_t n) {
                r.out.insert(r.out.end(), d, d + n);
                return 0;
            },
            &total, nullptr);
        r.status = st;
        if (st) r.msg = c->err;
    }
    std::lock_guard<std::mutex> g(S->mu);
    S->st.fallbacks += fallbacks;
}

void serve_worker(gzgpu_server *S, gzgpu_ctx *c) {
    for (;;) {
        std::vector<std::shared_ptr<ServeRequest>> batch;
        uint64_t reserve = 0, bid = 0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty(); });
            if (S->pending.empty()) return;  // stop requested and drained
            while (!S->pending.empty() && batch.size() < S->cfg.max_batch_requests) {
                auto &r = S->pending.front();
                // admit by free HBM; an oversized request runs alone
                if (!batch.empty() || S->hbm_in_use > 0)
                    if (S->hbm_in_use + reserve + r->reserve > S->cfg.hbm_budget_bytes) break;
                reserve += r->reserve;
                batch.push_back(r);
                S->pending.pop_front();
            }
            if (batch.empty()) {
                // budget held by batches in flight: wait for one to finish
                S->cv_done.wait(lk);
                continue;
            }
            S->hbm_in_use += reserve;
            S->st.hbm_peak_bytes = std::max<uint64_t>(S->st.hbm_peak_bytes, S->hbm_in_use);
            bid = S->next_batch++;
        }
        const uint64_t t0 = now_ns();
        for (auto &r : batch) {
            r->t_start = t0;
            r->batch = bid;
        }
        int st = guarded(c, [&]() {
            serve_batch(S, c, batch);
            return (int)GZGPU_OK;
        });
        const uint64_t t1 = now_ns();
        {
            std::lock_guard<std::mutex> g(S->mu);
            S->hbm_in_use -= reserve;
            S->st.batches++;
            for (auto &r : batch) {
                if (st != GZGPU_OK) {
                    r->status = st;
                    r->msg = c->err;
                }
                r->t_done = t1;
                S->st.requests_done++;
                S->st.bytes_in += r->in.size();
                S->st.bytes_out += r->out.size();
                std::vector<uint8_t>().swap(r->in);
            }
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

int gzgpu_server_create(int device, const gzgpu_server_config *cfg, gzgpu_server **out) {
    if (!out) return GZGPU_INVALID_ARGUMENT;
    *out = nullptr;
    gzgpu_server *S = new gzgpu_server();
    S->device = device;
    if (cfg) S->cfg = *cfg;
    if (!S->cfg.hbm_budget_bytes) S->cfg.hbm_budget_bytes = 200ull << 30;
    if (!S->cfg.max_batch_requests) S->cfg.max_batch_requests = 4096;
    if (!S->cfg.workers) S->cfg.workers = 2;
    for (uint32_t w = 0; w < S->cfg.workers; w++) {
        gzgpu_ctx *c = nullptr;
        int st = gzgpu_ctx_create(device, &c);
        if (st != GZGPU_OK) {
            for (auto *x : S->ctxs) gzgpu_ctx_destroy(x);
            delete S;
            return st;
        }
        S->ctxs.push_back(c);
    }
    for (auto *c : S->ctxs) S->workers.emplace_back(serve_worker, S, c);
    *out = S;
    return GZGPU_OK;
}

void gzgpu_server_destroy(gzgpu_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    S->cv_done.notify_all();
    for (auto &t : S->workers) t.join();
    for (auto *c : S->ctxs) gzgpu_ctx_destroy(c);
    delete S;
}

int gzgpu_submit(gzgpu_server *S, const uint8_t *in, size_t in_len, uint64_t *ticket) {
    if (!S || (!in && in_len) || !ticket) return GZGPU_INVALID_ARGUMENT;
    auto r = std::make_shared<ServeRequest>();
    r->in.assign(in, in + in_len);
    r->t_submit = now_ns();
    r->reserve = request_reserve(*r);
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->stop) return GZGPU_INVALID_ARGUMENT;
        r->id = S->next_id++;
        S->live[r->id] = r;
        S->pending.push_back(r);
        *ticket = r->id;
    }
    S->cv_work.notify_one();
    return GZGPU_OK;
}

int gzgpu_poll(gzgpu_server *S, uint64_t ticket, int32_t timeout_ms, gzgpu_request_info *info) {
    if (!S) return GZGPU_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->live.find(ticket);
    if (it == S->live.end()) return GZGPU_INVALID_ARGUMENT;
    std::shared_ptr<ServeRequest> r = it->second;
    auto done = [&] { return r->t_done != 0; };
    if (!done()) {
        if (timeout_ms == 0) return GZGPU_PENDING;
        if (timeout_ms < 0) S->cv_done.wait(lk, done);
        else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) return GZGPU_PENDING;
    }
    if (info) {
        info->status = r->status;
        info->members = r->members;
        info->out_len = r->out.size();
        info->queue_ms = (double)(r->t_start - r->t_submit) / 1e6;
        info->latency_ms = (double)(r->t_done - r->t_submit) / 1e6;
        info->batch = r->batch;
    }
    return GZGPU_OK;
}

int gzgpu_fetch(gzgpu_server *S, uint64_t ticket, uint8_t *out, size_t cap, size_t *out_len) {
    if (!S || !out_len) return GZGPU_INVALID_ARGUMENT;
    std::shared_ptr<ServeRequest> r;
    {
        std::lock_guard<std::mutex> g(S->mu);
        auto it = S->live.find(ticket);
        if (it == S->live.end()) return GZGPU_INVALID_ARGUMENT;
        r = it->second;
        if (!r->t_done) return GZGPU_PENDING;
        *out_len = r->out.size();
        if (r->out.size() > cap || (!out && !r->out.empty())) return GZGPU_OUTPUT_TOO_SMALL;
        S->live.erase(it);
    }
    if (!r->out.empty()) memcpy(out, r->out.data(), r->out.size());
    return r->status;
}

int gzgpu_server_get_stats(gzgpu_server *S, gzgpu_server_stats *st) {
    if (!S || !st) return GZGPU_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(S->mu);
    *st = S->st;
    return GZGPU_OK;
}

}  // extern "C"
