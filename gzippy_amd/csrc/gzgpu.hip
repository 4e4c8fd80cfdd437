// gzgpu.hip — host runtime and C ABI of the gfx950 parallel gzip decoder.
//
// Pipeline for a set of gzip members resident in HBM (one member = the
// single-member path, src/decompress/parallel/single_member.rs:987; many
// members = the member-per-worker path, src/decompress/bgzf.rs:2370):
//   1. envelope   RFC 1952 header / trailer per member (gzip_format.rs:47,135)
//   2. find       block-finder hits over large members' deflate bits
//                 (k_bf_stage_a / k_bf_stage_b)
//   3. plan       pieces on a chunk grid: piece 0 of each member starts at the
//                 deflate start with the all-zero window and writes straight
//                 into the member output; the others start at the first hit
//                 past each grid point and decode window-absent (markers)
//   4. inflate    one persistent launch, one wavefront per piece (k_inflate)
//   5. confirm    speculation check: piece j is accepted when piece j-1 ended
//                 exactly at its start (chunk_fetcher.rs:2506-2729 semantics);
//                 misses and slab overflows are re-decoded from the confirmed
//                 boundary in further launches
//   6. resolve    window chain + marker resolution (k_chain, k_resolve_body)
//   7. verify     CRC32 fold + GF(2) combine (k_crc_*) and ISIZE against the
//                 trailer (sm_driver.rs:212-224)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gzgpu.h"
#include "gz_common.h"
#include "k_blockfinder.h"
#include "k_crc.h"
#include "k_inflate.h"
#include "k_resolve.h"
#include "k_stored.h"

using namespace gz;

namespace {

constexpr uint64_t KIB = 1024, MIB = 1024 * 1024;
constexpr uint64_t NO_STOP = ~0ull;
constexpr size_t ENVELOPE_HEAD = 512;

struct HipError {
    hipError_t e;
    const char *what;
};

#define HIPCHK(x)                                              \
    do {                                                       \
        hipError_t e_ = (x);                                   \
        if (e_ != hipSuccess) throw HipError{e_, #x};          \
    } while (0)

struct StatusError {
    int status;
    std::string msg;
};

// Device memory of a call without allocation (gzgpu_decompress_device_noalloc):
// every buffer the call needs is carved from the device's `scratch` array,
// bump-allocated, nothing freed.  Set for the duration of such a call on the
// calling thread.
struct ScratchArena {
    uint8_t *base = nullptr;
    size_t cap = 0, used = 0;
};
thread_local ScratchArena *t_arena = nullptr;

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool borrowed = false;  // carved from a ScratchArena: never freed
    void ensure(size_t n) {
        if (n <= cap) return;
        if (t_arena) {
            const size_t c = (std::max<size_t>(n, 4096) + 255) & ~(size_t)255;
            if (t_arena->used + c > t_arena->cap)
                throw StatusError{GZGPU_OUT_OF_DEVICE_MEMORY,
                                  "no-allocation scratch exhausted: " + std::to_string(t_arena->used + c) + " of " +
                                      std::to_string(t_arena->cap) + " bytes needed"};
            p = t_arena->base + t_arena->used;
            t_arena->used += c;
            cap = c;
            borrowed = true;
            return;
        }
        if (p && !borrowed) HIPCHK(hipFree(p));
        p = nullptr;
        cap = 0;
        borrowed = false;
        size_t c = std::max<size_t>(n, 4096);
        hipError_t e = hipMalloc(&p, c);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw StatusError{GZGPU_OUT_OF_DEVICE_MEMORY, "hipMalloc(" + std::to_string(c) + ") failed"};
        }
        cap = c;
    }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        borrowed = false;
    }
    template <typename T>
    T *as() const {
        return reinterpret_cast<T *>(p);
    }
};

// ---------------------------------------------------------------- host CRC tables
struct CrcConst {
    uint32_t tab[8][256];
    uint32_t mx[4][256];  // follows tab: c -> c * x^(8 * 4096) mod P, one table per byte of c
    uint32_t xbyte[16][16];
    CrcConst() {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ CRC_POLY : c >> 1;
            tab[0][i] = c;
        }
        for (uint32_t i = 0; i < 256; i++)
            for (int t = 1; t < 8; t++) tab[t][i] = (tab[t - 1][i] >> 8) ^ tab[0][tab[t - 1][i] & 0xFF];
        // x^(2^k) for k = 0..; x^(8 * j * 16^k) = product over set bits
        uint32_t x2n[80];
        x2n[0] = 1u << 30;
        for (int k = 1; k < 80; k++) x2n[k] = mulmod(x2n[k - 1], x2n[k - 1]);
        for (int k = 0; k < 16; k++)
            for (int j = 0; j < 16; j++) {
                // exponent 8 * j * 16^k = j << (4k + 3)
                uint32_t p = 1u << 31;
                for (int b = 0; b < 4; b++)
                    if (j & (1 << b)) p = mulmod(x2n[4 * k + 3 + b], p);
                xbyte[k][j] = p;
            }
        {
            uint64_t nbytes = 4096;
            uint32_t X = 1u << 31;
            for (int k = 0; nbytes; k++, nbytes >>= 4)
                if (nbytes & 15) X = mulmod(xbyte[k][nbytes & 15], X);
            for (int k = 0; k < 4; k++)
                for (uint32_t b = 0; b < 256; b++) mx[k][b] = mulmod(X, b << (8 * k));
        }
    }
    static uint32_t mulmod(uint32_t a, uint32_t b) {
        uint32_t r = 0;
        for (uint32_t m = 1u << 31; m; m >>= 1) {
            if (a & m) r ^= b;
            b = (b & 1) ? (b >> 1) ^ CRC_POLY : b >> 1;
        }
        return r;
    }
};
const CrcConst &crc_const() {
    static CrcConst c;
    return c;
}


// RFC 1952 header length (gzip_format.rs:47-131).  0 = ok.
int parse_header(const uint8_t *p, size_t n, size_t *hdr_len) {
    if (n < 10) return GZGPU_INVALID_HEADER;
    if (p[0] != 0x1f || p[1] != 0x8b || p[2] != 8) return GZGPU_INVALID_HEADER;
    uint8_t flg = p[3];
    size_t off = 10;
    if (flg & 4) {
        if (off + 2 > n) return GZGPU_INVALID_HEADER;
        size_t xlen = (size_t)p[off] | (size_t)p[off + 1] << 8;
        off += 2;
        if (off + xlen > n) return GZGPU_INVALID_HEADER;
        off += xlen;
    }
    for (int f = 8; f <= 16; f <<= 1) {
        if (flg & f) {
            while (off < n && p[off]) off++;
            if (off >= n) return GZGPU_INVALID_HEADER;
            off++;
        }
    }
    if (flg & 2) {
        if (off + 2 > n) return GZGPU_INVALID_HEADER;
        off += 2;
    }
    *hdr_len = off;
    return GZGPU_OK;
}

uint32_t le32(const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

__global__ void k_gather_envelope(const uint8_t *__restrict__ in, const uint64_t *__restrict__ off,
                                  const uint64_t *__restrict__ len, uint32_t n, uint8_t *__restrict__ stage) {
    const uint32_t per = ENVELOPE_HEAD + 8;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)n * per;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t m = (uint32_t)(i / per), k = (uint32_t)(i % per);
        uint64_t L = len[m];
        uint8_t v = 0;
        if (k < ENVELOPE_HEAD) {
            if (k < L) v = in[off[m] + k];
        } else if (L >= 8) {
            v = in[off[m] + L - 8 + (k - ENVELOPE_HEAD)];
        }
        stage[i] = v;
    }
}

}  // namespace

// Pinned host staging for the small per-call uploads and downloads (piece
// descriptors, resolve lists, CRC segment lists, results).  A copy from
// pageable memory makes the host wait for the stream; from pinned memory it
// is a true async DMA, so the host keeps enqueueing while kernels run.  Bump
// allocation; reset() only after the stream has drained.
struct PinnedArena {
    std::vector<std::pair<uint8_t *, size_t>> blocks;
    size_t used = 0;
    // a fixed block instead (calls without allocation): bump only, no growth
    uint8_t *fixed = nullptr;
    size_t fixed_cap = 0, fixed_used = 0;
    void *alloc(size_t n) {
        n = (n + 63) & ~(size_t)63;
        if (fixed) {
            if (fixed_used + n > fixed_cap)
                throw StatusError{GZGPU_OUT_OF_DEVICE_MEMORY, "no-allocation host staging exhausted (" +
                                                                  std::to_string(fixed_cap) + " bytes)"};
            void *r = fixed + fixed_used;
            fixed_used += n;
            return r;
        }
        if (blocks.empty() || used + n > blocks.back().second) {
            size_t cap = std::max<size_t>(n, 1 << 20);
            if (!blocks.empty()) cap = std::max(cap, 2 * blocks.back().second);
            void *p = nullptr;
            if (hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                throw StatusError{GZGPU_GPU_UNAVAILABLE, "hipHostMalloc(" + std::to_string(cap) + ") failed"};
            }
            blocks.push_back({(uint8_t *)p, cap});
            used = 0;
        }
        void *r = blocks.back().first + used;
        used += n;
        return r;
    }
    // caller guarantees no copy from or to the arena is in flight
    void reset() {
        if (fixed) {
            fixed_used = 0;
            return;
        }
        if (blocks.size() > 1) {
            size_t want = 0;
            for (auto &b : blocks) want += b.second;
            release();
            alloc(want);
        }
        used = 0;
    }
    void release() {
        for (auto &b : blocks) (void)hipHostFree(b.first);
        blocks.clear();
        used = 0;
    }
};

// ---------------------------------------------------------------- context

struct Member {
    uint64_t ds = 0, de = 0;     // deflate bit range (absolute in d_in); de = trailer start
    uint64_t limit = 0;          // readable bits
    uint8_t *out = nullptr;
    uint64_t out_cap = 0;
    uint32_t exp_crc = 0, exp_isize = 0;
    bool split = false;
    bool stored = false;  // decoded by the stored-block split (no pieces)
    // results
    int status = GZGPU_OK;
    std::string msg;
    uint64_t out_len = 0;
    uint64_t end_bit = 0;
    uint32_t crc = 0;
    // the member's pieces: consecutive indices into the call's piece list
    // (no allocation per member)
    struct PieceRange {
        uint32_t first = 0, n = 0;
        size_t size() const { return n; }
        bool empty() const { return n == 0; }
        uint32_t operator[](size_t k) const { return first + (uint32_t)k; }
        uint32_t back() const { return first + n - 1; }
        void clear() { n = 0; }
        void push_back(uint32_t i) {
            if (n == 0) first = i;
            else if (i != first + n) throw StatusError{GZGPU_DECODE_FAILED, "member pieces not consecutive"};
            n++;
        }
    } pieces;  // in stream order
};

struct PieceHost {
    PieceDesc d{};
    uint64_t acc_lo = 0, acc_hi = 0;  // accepted predecessor end range
    uint32_t member = 0;
    bool confirmed = false, dropped = false;
    bool merged = false;  // folded into an earlier piece's re-decode: no output of its own
    uint64_t plan_cost = 0;  // planner's decode-cost estimate (0: from the member's average)
};

struct gzgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    // resident inflate wavefronts per CU, [ring u8 / u16][litlen root 9 / 10],
    // from the runtime's occupancy query (LDS and VGPRs, allocation granules)
    uint32_t inflate_occ[2][2] = {{16, 16}, {16, 16}};
    gzgpu_tuning tune{};
    std::string err;
    DevBuf d_lists;     // resolve and CRC lists of the current call, one upload
    DevBuf d_tile_crc;  // CRC32 of each resolve tile, from k_resolve_body
    DevBuf d_scan;      // end windows of the parallel-prefix window chain (two buffers)
    DevBuf d_in, d_out, d_pieces, d_pout, d_order, d_counters, d_surv, d_hits, d_slab, d_crc_tab, d_seg_off,
        d_seg_len, d_seg_crc, d_mseg, d_mlen, d_mcrc, d_env, d_env_off, d_env_len, d_bf_ranges, d_seg_tail;
    std::vector<DevBuf> extra_slabs;
    PinnedArena pin;
    // calls without allocation: the device's `scratch` array (its address on
    // this context's device) and a pinned staging block made at creation
    uint8_t *scratch = nullptr;
    uint8_t *noalloc_stage = nullptr;
    uint32_t max_pieces = 0;  // planner cap on pieces per call (0: none)
    // index builds: (bit after the member's deflate start, output offset) of
    // every block start the decode can resume at, collected by the pipeline
    std::vector<std::pair<uint64_t, uint64_t>> *cp_collect = nullptr;
    hipEvent_t ev[10]{};
    // host-buffer output: two pinned staging chunks, D2H of one while the
    // sink takes the other
    uint8_t *drain[2]{};
    hipEvent_t drain_ev[2]{};
    // streaming output (a single member to a sink): resolve in ranges, each
    // range's end event, D2H on a copy stream of their own while later ranges
    // resolve and the CRC runs
    hipStream_t copy_stream = nullptr;
    hipEvent_t res_ev[8]{};
    std::function<int(const uint8_t *, size_t)> *stream_sink = nullptr;
    bool streamed = false;
    int stream_rc = 0;
    gzgpu_stats st{};
};

namespace {
uint32_t lut_bits(const gzgpu_ctx *c) { return c->tune.lut_bits == 10 ? 10u : 9u; }
}  // namespace

namespace {

// Every device buffer of a context (destroy, and the swap around a call
// without allocation).
std::vector<DevBuf *> ctx_bufs(gzgpu_ctx *c) {
    return {&c->d_in,    &c->d_out,   &c->d_pieces,  &c->d_pout,      &c->d_order,   &c->d_counters, &c->d_surv,
            &c->d_hits,  &c->d_slab,  &c->d_crc_tab, &c->d_seg_off,   &c->d_seg_len, &c->d_seg_crc,  &c->d_mseg,
            &c->d_mlen,  &c->d_mcrc,  &c->d_env,     &c->d_env_off,   &c->d_env_len, &c->d_bf_ranges, &c->d_seg_tail,
            &c->d_lists, &c->d_scan,  &c->d_tile_crc};
}

void set_err(gzgpu_ctx *c, const std::string &m) {
    if (c) c->err = m;
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ms;
}

void upload(gzgpu_ctx *c, DevBuf &b, const void *src, size_t bytes) {
    b.ensure(bytes);
    if (!bytes) return;
    void *h = c->pin.alloc(bytes);
    memcpy(h, src, bytes);
    HIPCHK(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, c->stream));
}

// async device-to-host copy into pinned staging; valid after a stream sync
template <typename T>
T *download(gzgpu_ctx *c, const void *d, size_t count) {
    T *h = (T *)c->pin.alloc(count * sizeof(T));
    if (count) HIPCHK(hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    return h;
}

double host_ms() {
    return (double)std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
               .count() /
           1e3;
}

// GZGPU_DEBUG=1: per-launch timing lines on stderr; 2: also the planner's
// per-member split decisions
int debug_level() {
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("GZGPU_DEBUG");
        v = e && *e ? atoi(e) : 0;
    }
    return v;
}
bool debug_enabled() { return debug_level() >= 1; }

// The slicing tables and the x^(8 * 4096) multiply tables, contiguous on the
// device (uploaded once per context).
void ensure_crc_tables(gzgpu_ctx *c) {
    if (c->d_crc_tab.cap) return;
    const CrcConst &cc = crc_const();
    static_assert(offsetof(CrcConst, mx) == sizeof(CrcConst::tab), "tab, mx contiguous");
    c->d_crc_tab.ensure(sizeof(cc.tab) + sizeof(cc.mx));
    HIPCHK(hipMemcpy(c->d_crc_tab.p, cc.tab, sizeof(cc.tab) + sizeof(cc.mx), hipMemcpyHostToDevice));
}

// Indices of `keys` in ascending key order, equal keys by index (LSD radix,
// 11-bit digits over the bits the largest key uses): the host sorts a few
// thousand keys per call, where std::sort cost 0.1 ms.
std::vector<uint32_t> radix_order(const std::vector<uint64_t> &keys) {
    const size_t n = keys.size();
    std::vector<uint32_t> a(n), b(n);
    for (size_t i = 0; i < n; i++) a[i] = (uint32_t)i;
    uint64_t mx = 0;
    for (uint64_t k : keys) mx |= k;
    uint32_t cnt[2049];
    for (uint32_t sh = 0; sh < 64 && (mx >> sh); sh += 11) {
        memset(cnt, 0, sizeof(cnt));
        for (size_t i = 0; i < n; i++) cnt[((keys[a[i]] >> sh) & 2047) + 1]++;
        for (uint32_t d = 0; d < 2048; d++) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; i++) b[cnt[(keys[a[i]] >> sh) & 2047]++] = a[i];
        a.swap(b);
    }
    return a;
}

// ---- block finder over ranges [start_bit, end_bit) of d_in ----------------
// One stage-A launch over every range; hits of all ranges, sorted, duplicates
// (ranges that overlap) removed.  Every hit lies inside the range it came from.
// final_ok[i] (when given): range i also reports BFINAL = 1 dynamic headers
// (planning only; such hits carry bit HIT_FINAL_SHIFT of their info)
std::vector<gzgpu_boundary> find_hits(gzgpu_ctx *c, const uint8_t *d_in, uint64_t in_len,
                                      const std::vector<std::pair<uint64_t, uint64_t>> &rng, bool estimate = false,
                                      const std::vector<uint8_t> *final_ok = nullptr) {
    std::vector<gzgpu_boundary> out;
    std::vector<BfRange> table;
    uint64_t nchunk = 0, nbits = 0;
    for (size_t i = 0; i < rng.size(); i++) {
        const auto &r = rng[i];
        if (r.second <= r.first) continue;
        table.push_back({r.first, r.second, nchunk, final_ok && (*final_ok)[i] ? 1u : 0u, 0u});
        nchunk += (r.second - (r.first & ~31ull) + 31) / 32;
        nbits += r.second - r.first;
    }
    if (table.empty()) return out;
    uint32_t surv_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(nbits / 64, 4096), 1ull << 28);
    uint32_t hit_cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(nbits / 1024, 4096), 1ull << 26);
    upload(c, c->d_bf_ranges, table.data(), table.size() * sizeof(BfRange));
    for (int attempt = 0; attempt < 3; attempt++) {
        c->d_counters.ensure(64);
        c->d_surv.ensure((size_t)surv_cap * 8);
        c->d_hits.ensure((size_t)hit_cap * sizeof(gzgpu_boundary));
        HIPCHK(hipMemsetAsync(c->d_counters.p, 0, 64, c->stream));
        uint32_t *cnt = c->d_counters.as<uint32_t>();
        // stage A: one lane per 32 bit offsets (1 KiB of input per workgroup pass)
        uint64_t blocks = std::min<uint64_t>((nchunk + 255) / 256, (uint64_t)c->num_cus * 16);
        hipLaunchKernelGGL(k_bf_stage_a, dim3((uint32_t)blocks), dim3(256), 0, c->stream, d_in, in_len,
                           c->d_bf_ranges.as<BfRange>(), (uint32_t)table.size(), nchunk, c->d_surv.as<uint64_t>(),
                           surv_cap, cnt + 0, c->d_hits.as<gzgpu_boundary>(), hit_cap, cnt + 1);
        HIPCHK(hipGetLastError());
        c->st.launches++;
        // stage B reads the survivor count on the device; one read-back of
        // both counters and the first hits (pinned), the rest only if needed
        const uint32_t bb = std::min<uint32_t>(surv_cap / 256 + 1, (uint32_t)c->num_cus * 16);
        hipLaunchKernelGGL(k_bf_stage_b, dim3(bb), dim3(256), 0, c->stream, d_in, in_len, c->d_surv.as<uint64_t>(),
                           cnt + 0, surv_cap, c->d_hits.as<gzgpu_boundary>(), hit_cap, cnt + 1);
        HIPCHK(hipGetLastError());
        c->st.launches++;
        if (estimate) {
            const uint32_t be = std::min<uint32_t>(hit_cap, (uint32_t)c->num_cus * 64);  // one wavefront per hit in flight, several per CU
            hipLaunchKernelGGL(k_bf_estimate, dim3(be), dim3(64), 0, c->stream, d_in, in_len,
                               c->d_hits.as<gzgpu_boundary>(), cnt + 1, hit_cap);
            HIPCHK(hipGetLastError());
            c->st.launches++;
        }
        const uint32_t first = std::min<uint32_t>(hit_cap, 8192);
        const uint32_t *h_cnt = download<uint32_t>(c, cnt, 2);
        const gzgpu_boundary *h_hits = download<gzgpu_boundary>(c, c->d_hits.p, first);
        HIPCHK(hipStreamSynchronize(c->stream));
        if (h_cnt[0] > surv_cap || h_cnt[1] > hit_cap) {
            if (h_cnt[0] > surv_cap) surv_cap = h_cnt[0] + 1024;
            if (h_cnt[1] > hit_cap) hit_cap = h_cnt[1] + 1024;
            continue;
        }
        const double ts = debug_enabled() ? host_ms() : 0;
        const uint32_t nh = h_cnt[1];
        out.assign(h_hits, h_hits + std::min(nh, first));
        if (nh > first) {
            const gzgpu_boundary *rest = download<gzgpu_boundary>(c, c->d_hits.as<gzgpu_boundary>() + first, nh - first);
            HIPCHK(hipStreamSynchronize(c->stream));
            out.insert(out.end(), rest, rest + (nh - first));
        }
        // by (bit_offset, seek_bit): radix on bit_offset, then the rare runs
        // of equal bit_offset by seek_bit
        {
            std::vector<uint64_t> keys(out.size());
            for (size_t i = 0; i < out.size(); i++) keys[i] = out[i].bit_offset;
            const std::vector<uint32_t> ord = radix_order(keys);
            std::vector<gzgpu_boundary> sorted(out.size());
            for (size_t i = 0; i < ord.size(); i++) sorted[i] = out[ord[i]];
            out.swap(sorted);
            for (size_t i = 0; i < out.size();) {
                size_t j = i + 1;
                while (j < out.size() && out[j].bit_offset == out[i].bit_offset) j++;
                if (j - i > 1)
                    std::sort(out.begin() + i, out.begin() + j,
                              [](const gzgpu_boundary &x, const gzgpu_boundary &y) { return x.seek_bit < y.seek_bit; });
                i = j;
            }
        }
        if (table.size() > 1)
            out.erase(std::unique(out.begin(), out.end(),
                                  [](const gzgpu_boundary &a, const gzgpu_boundary &b) {
                                      return a.bit_offset == b.bit_offset && a.seek_bit == b.seek_bit;
                                  }),
                      out.end());
        if (debug_enabled())
            fprintf(stderr, "[gzgpu] host: hits read-back + sort %.3f ms (%u stage-A survivors, %u hits, %llu bits)\n",
                    host_ms() - ts, h_cnt[0], nh, (unsigned long long)nbits);
        return out;
    }
    throw StatusError{GZGPU_OUT_OF_DEVICE_MEMORY, "block finder: candidate buffers kept overflowing"};
}

std::vector<gzgpu_boundary> find_hits(gzgpu_ctx *c, const uint8_t *d_in, uint64_t in_len, uint64_t start_bit,
                                      uint64_t end_bit) {
    return find_hits(c, d_in, in_len, std::vector<std::pair<uint64_t, uint64_t>>{{start_bit, end_bit}});
}

uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// The litlen LUT root width of the inflate kernel (tuning lut_bits: 0 = 9).
uint32_t lut_bits(const gzgpu_ctx *c);
// Resident inflate wavefronts per CU for a ring type and LUT root width.
template <typename RT>
uint32_t inflate_waves_per_cu(const gzgpu_ctx *c, uint32_t lw) {
    return c->inflate_occ[sizeof(RT) == 2 ? 1 : 0][lw == 10 ? 1 : 0];
}
// planner: smallest piece (compressed KiB; the reference's block finder
// keeps 32 KiB, gzip_block_finder.rs:109) and pieces per resident wavefront.
// With starts inside blocks a member's pieces come out equal, so one launch
// round with one piece per resident wavefront (less a 1/32 margin: a piece
// more than the slots waits a whole piece time) beats several: every piece
// also pays a marker (u16) phase of its own (~160k entries on text, each
// ~20% slower than after the flip), and the window chain and marker
// resolution grow with the piece count.  A member whose cost per compressed bit is high
// (member_cost) may go down to MIN_SPACING_FLOOR_KIB so that no piece costs
// much more than MIN_PIECE_OUT.
constexpr uint64_t MIN_SPACING_KIB = 2, MIN_SPACING_FLOOR_KIB = 1;
constexpr uint64_t MIN_PIECE_OUT = 64 * 1024;
constexpr double PIECES_PER_SLOT = 1.0, SLOT_FRACTION = 31.0 / 32.0;
// a start inside a block: at least MID_HEAD bits after the block's header
// hit (a dynamic header is rarely longer; a start inside it fails its sync
// and costs a re-decode) and MID_TAIL bits before the next hit
constexpr uint64_t MID_HEAD = 2048, MID_TAIL = 2048;
// The last FINAL_SCAN_BITS of every member to be split are scanned for
// BFINAL = 1 dynamic headers too: the final block then takes starts inside
// it like any other (the reference finder, and the public find_blocks, report
// BFINAL = 0 only, so without this the piece reaching the final block decodes
// all of it).  512 KiB covers the final block of the encoders measured; a
// longer one keeps the old behaviour (its piece runs to the end).
constexpr uint64_t FINAL_SCAN_BITS = 8ull * 512 * 1024;
// planner: piece-size slope along a member (see the grid in run_members_aligned)
constexpr double TILT = 0.25;
// speculation: from this launch on a miss re-decodes the member's whole
// remainder as one piece; a member still unfinished after SPEC_MAX_ROUNDS
// launches fails alone (cannot happen for a valid stream: the remainder
// piece either reaches the final block or meets a real error)
constexpr int SPEC_FALLBACK_ROUND = 4, SPEC_MAX_ROUNDS = 16;
// planner: members cut into at least this many pieces get per-block output
// estimates from the block finder
constexpr uint64_t ESTIMATE_MIN_PIECES = 64;

// Decode cost of a member in output-entry units.  A wavefront's decode time
// at full occupancy is about a per compressed bit plus b per output entry:
// a round covers at most 128 input bits, and a round's copies take time in
// proportion to their bytes.  Fitted on per-piece kernel times of the
// inflate launch (1 GiB text member: 83k bits, 270k entries, 3.23 ms;
// 256 MiB of u32 deltas: 182k bits, 68k entries, 4.58 ms): a = 23 ns,
// b = 4.7 ns, so a Huffman-coded member costs 5 x bits + ISIZE.  Stored
// data (ratio below 1.15, the reference's StoredParallel cut,
// decompress/mod.rs:143-151) moves at copy speed, ~440 MB/s for a wavefront
// among fifteen others, and is weighted 1/16 per bit.  (The earlier
// max(4 x bits, ISIZE) with stored data at 1/32 per bit, fitted on 64-bit
// rounds, planned the hipGraph batch 7% slower.)  ISIZE is the size mod 2^32;
// a member whose true size wrapped looks stored-like here, which only
// changes how it is split.
uint64_t member_cost(uint64_t cbits, uint32_t exp_isize) {
    const bool stored_like = (uint64_t)exp_isize * 100 < (cbits / 8) * 115;
    return std::max<uint64_t>(1, stored_like ? cbits / 16 : 5 * cbits + exp_isize);
}

// Entries per compressed bit for each 12-bit output estimate of the block
// finder (log2 in 1/256 steps, offset 8).
const double *ratio_table() {
    static const std::vector<double> t = [] {
        std::vector<double> v(4096);
        for (int q = 0; q < 4096; q++) v[q] = std::exp2((double)q / 256.0 - 8.0);
        return v;
    }();
    return t.data();
}

uint64_t marker_slab_bytes(uint64_t expect_entries) { return round_up(2 * (expect_entries + 64 * KIB) + 512, 256); }

// Starts inside blocks (MIDBLOCK pieces); GZGPU_MIDBLOCK=0 turns them off.
// Chains longer than this resolve their end windows as a parallel prefix
// (k_chain_scan_*) instead of piece by piece (k_chain).  k_chain costs about
// 17 us per piece of a chain; a prefix step about 0.1 ms over all slots.
// GZGPU_CHAIN_SCAN_MIN overrides it (0: every chain, a large value: none).
uint32_t chain_scan_min() {
    const char *e = getenv("GZGPU_CHAIN_SCAN_MIN");
    return e && *e ? (uint32_t)strtoul(e, nullptr, 10) : 16u;
}

bool midblock_enabled() {
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("GZGPU_MIDBLOCK");
        v = !(e && *e == '0');
    }
    return v == 1;
}

// Issue-priority balance inside the inflate launch (progress_report);
// GZGPU_PROG=0 turns it off.
bool prog_balance() {
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("GZGPU_PROG");
        v = !(e && *e == '0');
    }
    return v == 1;
}

// One launch of the persistent inflate kernel over `descs` (pulled
// longest-first by `cost`) on the context stream; finish_inflate waits for
// it and reads the per-piece results back.
struct InflateLaunch {
    const PieceOut *hro = nullptr;
    uint32_t n = 0, grid = 0;
    bool u16 = false;
};

template <typename RT>
InflateLaunch enqueue_inflate(gzgpu_ctx *c, const uint8_t *d_in, uint64_t in_len, const std::vector<PieceDesc> &descs,
                              const std::vector<uint64_t> &cost) {
    const double t0 = debug_enabled() ? host_ms() : 0;
    InflateLaunch L;
    L.n = (uint32_t)descs.size();
    L.u16 = sizeof(RT) == 2;
    // longest first: radix order of the complement of each cost's top 16
    // significant bits (a 6-bit exponent and 10-bit mantissa; LPT needs no
    // finer order), equal keys by index
    std::vector<uint64_t> keys(L.n);
    for (uint32_t i = 0; i < L.n; i++) {
        const uint64_t v = cost[i];
        const uint32_t e = v ? 63u - (uint32_t)__builtin_clzll(v) : 0u;
        const uint64_t q = e < 10 ? v : ((uint64_t)(e - 9) << 10) | ((v >> (e - 10)) & 1023);
        keys[i] = 0xFFFFu - (q & 0xFFFFu);
    }
    const std::vector<uint32_t> order = radix_order(keys);
    const double t1 = debug_enabled() ? host_ms() : 0;
    upload(c, c->d_pieces, descs.data(), descs.size() * sizeof(PieceDesc));
    upload(c, c->d_order, order.data(), order.size() * 4);
    if (debug_enabled()) fprintf(stderr, "[gzgpu] host: sort %.3f ms upload %.3f ms\n", t1 - t0, host_ms() - t1);
    c->d_pout.ensure(descs.size() * sizeof(PieceOut));
    c->d_counters.ensure(64);
    HIPCHK(hipMemsetAsync(c->d_counters.p, 0, 64, c->stream));
    // LDS bounds residency (160 KiB per CU): 10 KB per wavefront
    const uint32_t lw = lut_bits(c);
    uint32_t per_cu = inflate_waves_per_cu<RT>(c, lw);
    if (c->tune.waves_per_cu) per_cu = std::min<uint32_t>(per_cu, c->tune.waves_per_cu);
    L.grid = (uint32_t)std::min<uint64_t>(L.n, (uint64_t)c->num_cus * per_cu);
    // issue-priority balance when every piece is resident from the start
    // (progress fractions are then comparable): the launch-wide progress
    // counter (zeroed with the queue; each piece adds up to PROG_UNIT)
    uint64_t *prog = nullptr;
    float prog_inv = 0.f;
    if (prog_balance() && L.n <= L.grid) {
        prog = (uint64_t *)(c->d_counters.as<uint8_t>() + 8);
        prog_inv = 1.0f / ((float)L.n * PROG_UNIT);
    }
    if (debug_enabled()) fprintf(stderr, "[gzgpu] host: inflate setup %.3f ms\n", host_ms() - t0);
    HIPCHK(hipEventRecord(c->ev[5], c->stream));
    if (lw == 10)
        hipLaunchKernelGGL((k_inflate<RT, 10>), dim3(L.grid), dim3(64), 0, c->stream, d_in, in_len,
                           c->d_pieces.as<PieceDesc>(), c->d_order.as<uint32_t>(), L.n, c->d_counters.as<uint32_t>(),
                           c->d_pout.as<PieceOut>(), (const uint32_t *)nullptr, prog, prog_inv);
    else
        hipLaunchKernelGGL((k_inflate<RT, 9>), dim3(L.grid), dim3(64), 0, c->stream, d_in, in_len,
                           c->d_pieces.as<PieceDesc>(), c->d_order.as<uint32_t>(), L.n, c->d_counters.as<uint32_t>(),
                           c->d_pout.as<PieceOut>(), (const uint32_t *)nullptr, prog, prog_inv);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[6], c->stream));
    c->st.launches++;
    L.hro = download<PieceOut>(c, c->d_pout.p, L.n);
    return L;
}

std::vector<PieceOut> finish_inflate(gzgpu_ctx *c, const InflateLaunch &L, const std::vector<PieceDesc> &descs,
                                     float &ms) {
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<PieceOut> ro(L.hro, L.hro + L.n);
    ms = ev_ms(c->ev[5], c->ev[6]);
    c->st.ms_inflate_kernel += ms;
    c->st.inflate_launches++;
    if (debug_enabled()) {
        std::vector<double> us;
        uint64_t t0 = ~0ull, t1 = 0, bytes = 0, flipped = 0, n16 = 0;
        for (auto &r : ro) {
            us.push_back((double)(r.t_end - r.t_begin) / 100.0);
            t0 = std::min(t0, r.t_begin);
            t1 = std::max(t1, r.t_end);
            bytes += r.n16 + r.n8;
            n16 += r.n16;
            flipped += r.mode == 2;
        }
        size_t slow = 0;
        for (size_t i = 1; i < ro.size(); i++)
            if (ro[i].t_end - ro[i].t_begin > ro[slow].t_end - ro[slow].t_begin) slow = i;
        if (debug_level() >= 3)
            for (size_t i = 0; i < ro.size(); i++)
                fprintf(stderr, "[gzgpu] piece %zu n16 %llu n8 %llu bits %llu us %.1f begin %.1f flags %u\n", i,
                        (unsigned long long)ro[i].n16, (unsigned long long)ro[i].n8,
                        (unsigned long long)ro[i].bits_done, (double)(ro[i].t_end - ro[i].t_begin) / 100.0,
                        (double)(ro[i].t_begin - t0) / 100.0, descs[i].flags);
        std::sort(us.begin(), us.end());
        fprintf(stderr,
                "[gzgpu] inflate<%s> pieces=%u grid=%u kernel=%.3f ms t0=%.1f us span=%.3f ms piece_us min=%.1f p50=%.1f "
                "max=%.1f entries=%llu (u16 %llu) flipped=%llu per-wave MB/s p50=%.1f | slowest: %llu entries, "
                "%llu bits, %u blocks, flags %u, status %u\n",
                L.u16 ? "u16" : "u8", L.n, L.grid, ms, (double)(t0 % 100000000ull) / 100.0, (double)(t1 - t0) / 1e5, us.front(), us[us.size() / 2],
                us.back(), (unsigned long long)bytes, (unsigned long long)n16, (unsigned long long)flipped,
                us[us.size() / 2] > 0 ? (double)bytes / L.n / us[us.size() / 2] : 0.0,
                (unsigned long long)(ro[slow].n16 + ro[slow].n8), (unsigned long long)ro[slow].bits_done,
                ro[slow].nblocks, descs[slow].flags, ro[slow].status);
    }
    return ro;
}

// Inflate the pieces in `todo` in one persistent launch; results land in R.
// A call with marker pieces runs all its pieces in the u16-ring kernel (a
// marker piece flips to u8 output in place), one without in the u8-ring
// kernel.  Longest first.
void launch_inflate(gzgpu_ctx *c, const uint8_t *d_in, uint64_t in_len, std::vector<PieceHost> &P,
                    const std::vector<Member> &mem, std::vector<uint32_t> &todo, std::vector<PieceOut> &R) {
    auto stop_of = [&](const PieceDesc &d) { return d.stop_bit == NO_STOP ? mem[d.member].de : d.stop_bit; };
    if (todo.empty()) return;
    const double tl = debug_enabled() ? host_ms() : 0;
    std::vector<PieceDesc> d;
    std::vector<uint64_t> cost;  // the member's cost per compressed bit x the piece's bits
    bool any_marker = false;
    for (uint32_t pi : todo) {
        d.push_back(P[pi].d);
        const Member &M = mem[P[pi].d.member];
        const uint64_t mbits = std::max<uint64_t>(1, M.de - M.ds);
        const double per_bit = (double)member_cost(mbits, M.exp_isize) / (double)mbits;
        const uint64_t bits = stop_of(P[pi].d) - P[pi].d.start_bit;
        cost.push_back(P[pi].plan_cost ? P[pi].plan_cost : (uint64_t)(per_bit * 256.0 * (double)bits));
        any_marker |= (P[pi].d.flags & PF_MARKER) != 0;
    }
    // issue-priority weights for a launch with more pieces than wavefronts
    // (progress_report); with every piece resident the launch-wide progress
    // counter is used instead and the weights are ignored
    if (prog_balance()) {
        uint64_t mx = 1;
        for (uint64_t v : cost) mx = std::max(mx, v);
        for (size_t i = 0; i < d.size(); i++) d[i].weight = (float)((double)cost[i] / (double)mx);
    }
    float ms = 0;
    if (debug_enabled()) fprintf(stderr, "[gzgpu] host: launch lists %.3f ms\n", host_ms() - tl);
    std::vector<PieceOut> r = any_marker ? finish_inflate(c, enqueue_inflate<uint16_t>(c, d_in, in_len, d, cost), d, ms)
                                         : finish_inflate(c, enqueue_inflate<uint8_t>(c, d_in, in_len, d, cost), d, ms);
    uint64_t in_bytes = 0, out_bytes = 0;
    for (size_t i = 0; i < todo.size(); i++) {
        R[todo[i]] = r[i];
        in_bytes += (r[i].end_bit > d[i].start_bit ? r[i].end_bit - d[i].start_bit : 0) / 8;
        out_bytes += 2 * r[i].n16 + r[i].n8;
    }
    if (any_marker) {
        c->st.ms_inflate_marker += ms;
        c->st.marker_in_bytes += in_bytes;
        c->st.marker_out_bytes += out_bytes;
    } else {
        c->st.ms_inflate_clean += ms;
        c->st.clean_in_bytes += in_bytes;
        c->st.clean_out_bytes += out_bytes;
    }
}

// Every kernel reads the compressed input from a 16-byte-aligned base: the
// inflate kernels fetch it with scalar loads, which ignore the low address
// bits.  An unaligned caller pointer becomes (base, shift bytes) and every
// bit position is moved by the shift.
struct AlignedIn {
    const uint8_t *base;
    uint64_t shift;  // bytes
};
AlignedIn align_input(const void *p) {
    const uintptr_t a = (uintptr_t)p;
    return {(const uint8_t *)(a & ~(uintptr_t)15), (uint64_t)(a & 15)};
}

void run_members_aligned(gzgpu_ctx *c, const uint8_t *d_in, uint64_t in_len, std::vector<Member> &mem);

// ---- the member pipeline ---------------------------------------------------
void run_members(gzgpu_ctx *c, const uint8_t *d_in, uint64_t in_len, std::vector<Member> &mem) {
    const AlignedIn ai = align_input(d_in);
    const uint64_t sb = ai.shift * 8;
    for (auto &M : mem) {
        M.ds += sb;
        M.de += sb;
        M.limit += sb;
    }
    run_members_aligned(c, ai.base, in_len + ai.shift, mem);
    for (auto &M : mem) {
        M.ds -= sb;
        M.de -= sb;
        M.limit -= sb;
        if (M.end_bit >= sb) M.end_bit -= sb;
    }
}

// Streaming output: a single member of at least STREAM_MIN bytes to a sink
// resolves in STREAM_RANGES launches, and the D2H of each finished range
// overlaps the later ones.
constexpr uint32_t STREAM_RANGES = 8;
constexpr uint64_t STREAM_MIN = 16 * MIB;
int drain_ranges(gzgpu_ctx *c, const uint8_t *d, uint64_t n, const std::vector<std::pair<uint64_t, int>> &ready);

// StoredParallel (decompress_stored_parallel, stored_split.rs:265; routed by
// parallel_sm_unprofitable && first_block_is_stored, decompress/mod.rs:282-286):
// a member whose output is no larger than 1.15x its compressed bytes and
// which is one chain of stored blocks from its first byte to its trailer is
// copied block by block, with no block finder and no speculation.  The chain
// is walked on the host through the byte-aligned LEN/NLEN candidates of
// k_stored_scan; a member whose chain breaks (a Huffman block, or a final
// block that ends before the trailer) goes through the speculative path, as
// the reference's NotStoredDominated does.  Only members of at least
// STORED_MIN_BYTES: a smaller one is one piece anyway.
constexpr uint64_t STORED_MIN_BYTES = 1 * MIB;

void stored_members(gzgpu_ctx *c, const uint8_t *d_in, uint64_t in_len, std::vector<Member> &mem) {
    std::vector<uint32_t> cand_m;
    for (uint32_t m = 0; m < mem.size(); m++) {
        const Member &M = mem[m];
        const uint64_t cbytes = (M.de - M.ds) / 8;
        if (M.status != GZGPU_OK || !M.split || cbytes < STORED_MIN_BYTES || M.ds % 8 || M.de % 8) continue;
        if ((uint64_t)M.exp_isize * 100 >= cbytes * 115) continue;  // parallel_sm_unprofitable
        cand_m.push_back(m);
    }
    if (cand_m.empty()) return;
    // one scan per member into one candidate list (a few per MiB besides the
    // true headers: LEN/NLEN coincidences)
    uint64_t cap = 0;
    for (uint32_t m : cand_m) cap += (mem[m].de - mem[m].ds) / 8 / 8192 + 4096;
    cap = std::min<uint64_t>(cap, 1ull << 28);
    for (int attempt = 0; attempt < 2; attempt++) {
        c->d_counters.ensure(64);
        c->d_surv.ensure(cap * 8);
        HIPCHK(hipMemsetAsync(c->d_counters.p, 0, 64, c->stream));
        for (uint32_t m : cand_m) {
            const uint64_t first = mem[m].ds / 8, last = mem[m].de / 8;
            const uint32_t blocks = (uint32_t)std::min<uint64_t>((last - first + 1023) / 1024 + 1, (uint64_t)c->num_cus * 8);
            hipLaunchKernelGGL(k_stored_scan, dim3(blocks), dim3(256), 0, c->stream, d_in, in_len, first, last,
                               c->d_surv.as<unsigned long long>(), (uint32_t)cap, c->d_counters.as<uint32_t>());
            HIPCHK(hipGetLastError());
            c->st.launches++;
        }
        const uint32_t *h_cnt = download<uint32_t>(c, c->d_counters.p, 1);
        HIPCHK(hipStreamSynchronize(c->stream));
        const uint32_t nc = h_cnt[0];
        if (nc > cap) {
            cap = nc + 1024;
            continue;
        }
        const uint64_t *hc = download<uint64_t>(c, c->d_surv.p, nc);
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<uint64_t> keys(nc);
        for (uint32_t i = 0; i < nc; i++) keys[i] = hc[i] & ((1ull << 40) - 1);
        const std::vector<uint32_t> ord = radix_order(keys);
        std::vector<StoredRun> runs;
        for (uint32_t m : cand_m) {
            Member &M = mem[m];
            const uint64_t first = M.ds / 8, last = M.de / 8;
            size_t k = (size_t)(std::lower_bound(ord.begin(), ord.end(), first,
                                                 [&](uint32_t i, uint64_t v) { return keys[i] < v; }) -
                                ord.begin());
            uint64_t p = first, out = 0;
            const size_t run0 = runs.size();
            bool ok = false;
            for (;;) {
                while (k < ord.size() && keys[ord[k]] < p) k++;
                if (k == ord.size() || keys[ord[k]] != p) break;  // not a stored header: the chain ends
                const uint64_t v = hc[ord[k]];
                const uint32_t len = (uint32_t)(v >> 40) & 0xFFFF;
                if (c->cp_collect) c->cp_collect->push_back({p * 8 - M.ds, out});
                if (len && out + len <= M.out_cap) runs.push_back({p + 5, (uint64_t)(uintptr_t)(M.out + out), len, 0});
                out += len;
                p += 5 + len;
                if (v >> 56) {
                    ok = p == last;  // the trailer right after the final block
                    break;
                }
                if (p >= last) break;
            }
            if (!ok || out > M.out_cap || out != M.exp_isize) {
                // not pure-stored (or a size the trailer contradicts): the
                // speculative path decodes it and reports what it finds
                runs.resize(run0);
                continue;
            }
            M.stored = true;
            M.out_len = out;
            M.end_bit = last * 8;
            c->st.stored_members++;
            c->st.stored_blocks += (uint32_t)(runs.size() - run0);
        }
        if (runs.empty()) return;
        upload(c, c->d_order, runs.data(), runs.size() * sizeof(StoredRun));
        hipLaunchKernelGGL(k_stored_copy, dim3((uint32_t)std::min<size_t>(runs.size(), (size_t)c->num_cus * 8)),
                           dim3(256), 0, c->stream, d_in, in_len, c->d_order.as<StoredRun>(), (uint32_t)runs.size());
        HIPCHK(hipGetLastError());
        c->st.launches++;
        return;
    }
}

// Stored runs inside members to be split.  A chain of STORED_RUN_MIN or more
// stored blocks whose byte-aligned headers follow each other (q -> q + 5 +
// LEN, each LEN == ~NLEN: a mixed member's random segments) holds no block
// header of this stream in its payloads, so the bit-level block finder skips
// them: on a 2 GiB mixed member they are two thirds of the compressed bits.
// (The chain's first block, whose header need not be byte-aligned, is found
// by the finder as before.)  Only for members whose output is less than
// STORED_SCAN_MAX_RATIO times their compressed bytes: better-compressed
// members hold few stored bytes and skip the scan and its read-back.
// Returns, per member, its payload bit ranges to skip, sorted.
constexpr uint32_t STORED_RUN_MIN = 3;
constexpr uint64_t STORED_SCAN_MAX_RATIO = 4, STORED_SCAN_MIN_BYTES = 4 * MIB;
constexpr uint64_t STORED_SKIP_MARGIN = 64;  // payload bits kept at either end of a skipped range

std::vector<std::vector<std::pair<uint64_t, uint64_t>>> stored_payloads(gzgpu_ctx *c, const uint8_t *d_in,
                                                                        uint64_t in_len, const std::vector<Member> &mem,
                                                                        const std::vector<uint8_t> &want) {
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> out(mem.size());
    std::vector<uint32_t> ms;
    uint64_t cap = 4096;
    for (uint32_t m = 0; m < mem.size(); m++) {
        const Member &M = mem[m];
        const uint64_t cbytes = (M.de - M.ds) / 8;
        if (!want[m] || cbytes < STORED_SCAN_MIN_BYTES || (uint64_t)M.exp_isize >= STORED_SCAN_MAX_RATIO * cbytes)
            continue;
        ms.push_back(m);
        cap += cbytes / 1024;
    }
    if (ms.empty()) return out;
    cap = std::min<uint64_t>(cap, 1ull << 28);
    for (int attempt = 0; attempt < 2; attempt++) {
        c->d_counters.ensure(64);
        c->d_surv.ensure(cap * 8);
        HIPCHK(hipMemsetAsync(c->d_counters.p, 0, 64, c->stream));
        for (uint32_t m : ms) {
            const uint64_t first = (mem[m].ds + 7) / 8, last = mem[m].de / 8;
            const uint32_t blocks = (uint32_t)std::min<uint64_t>((last - first + 1023) / 1024 + 1, (uint64_t)c->num_cus * 8);
            hipLaunchKernelGGL(k_stored_scan, dim3(blocks), dim3(256), 0, c->stream, d_in, in_len, first, last,
                               c->d_surv.as<unsigned long long>(), (uint32_t)cap, c->d_counters.as<uint32_t>());
            HIPCHK(hipGetLastError());
            c->st.launches++;
        }
        const uint32_t *h_cnt = download<uint32_t>(c, c->d_counters.p, 1);
        HIPCHK(hipStreamSynchronize(c->stream));
        const uint32_t nc = h_cnt[0];
        if (nc > cap) {
            cap = (uint64_t)nc + 1024;
            continue;
        }
        const uint64_t *hc = download<uint64_t>(c, c->d_surv.p, nc);
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<uint64_t> keys(nc);
        for (uint32_t i = 0; i < nc; i++) keys[i] = hc[i] & ((1ull << 40) - 1);
        const std::vector<uint32_t> ord = radix_order(keys);
        std::vector<uint64_t> pos(nc);  // candidate byte offsets, ascending
        for (uint32_t i = 0; i < nc; i++) pos[i] = keys[ord[i]];
        std::vector<uint8_t> seen(nc, 0);
        // candidate at byte q, searched forward from index `from` (a chain's
        // next header follows the current one, usually as the very next
        // candidate) in doubling steps, then by bisection inside the last step
        auto find = [&](size_t from, uint64_t q) -> int64_t {
            size_t lo = from, hi = from, step = 1;
            while (hi < nc && pos[hi] < q) {
                lo = hi + 1;
                hi += step;
                step <<= 1;
            }
            hi = std::min<size_t>(hi, nc);  // every candidate before lo is below q, pos[hi] is not
            const size_t at = (size_t)(std::lower_bound(pos.begin() + lo, pos.begin() + hi, q) - pos.begin());
            return at < nc && pos[at] == q ? (int64_t)at : -1;
        };
        for (uint32_t m : ms) {
            const Member &M = mem[m];
            const uint64_t first = (M.ds + 7) / 8, last = M.de / 8;
            size_t k = (size_t)(std::lower_bound(pos.begin(), pos.end(), first) - pos.begin());
            for (; k < nc && pos[k] < last; k++) {
                if (seen[k]) continue;
                // follow the chain from candidate k
                std::vector<std::pair<uint64_t, uint32_t>> chain;  // (header byte, LEN)
                int64_t j = (int64_t)k;
                while (j >= 0 && !seen[j]) {
                    seen[j] = 1;
                    const uint64_t v = hc[ord[j]];
                    const uint32_t len = (uint32_t)(v >> 40) & 0xFFFF;
                    chain.push_back({pos[j], len});
                    if (v >> 56) break;  // BFINAL
                    j = find((size_t)j + 1, pos[j] + 5 + len);
                }
                if (chain.size() < STORED_RUN_MIN) continue;
                for (auto &b : chain) {
                    const uint64_t p0 = (b.first + 5) * 8 + STORED_SKIP_MARGIN, p1 = (b.first + 5 + b.second) * 8;
                    if (p1 > p0 + STORED_SKIP_MARGIN) out[m].push_back({p0, p1 - STORED_SKIP_MARGIN});
                }
            }
            std::sort(out[m].begin(), out[m].end());
        }
        return out;
    }
    return out;  // candidate lists kept overflowing: nothing skipped
}

void run_members_aligned(gzgpu_ctx *c, const uint8_t *d_in, uint64_t in_len, std::vector<Member> &mem) {
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    stored_members(c, d_in, in_len, mem);
    const uint64_t chunk_bits = (c->tune.chunk_bytes ? c->tune.chunk_bytes : 512 * KIB) * 8;
    const uint64_t min_spacing = 8 * MIN_SPACING_KIB * KIB;
    auto min_spacing_of = [&](const Member &M) -> uint64_t {
        const uint64_t cbits = M.de - M.ds;
        const double bits = (double)MIN_PIECE_OUT * (double)cbits /
                            (double)member_cost(cbits, M.exp_isize);
        return std::min<uint64_t>(min_spacing, std::max<uint64_t>(8 * MIN_SPACING_FLOOR_KIB * KIB, (uint64_t)bits));
    };
    const uint32_t lut_waves = inflate_waves_per_cu<uint16_t>(c, lut_bits(c));
    const uint64_t slots = (uint64_t)c->num_cus *  // resident inflate wavefronts
                           (c->tune.waves_per_cu ? std::min<uint32_t>(c->tune.waves_per_cu, lut_waves) : lut_waves);
    // Piece grid.  Explicit pieces_per_chunk: chunk_bits / ppc compressed bits.
    // Default: balance decode time (member_cost); every piece aims at one
    // resident slot's share of the call's total cost, and a member's pieces
    // split its compressed bits evenly.
    auto cost_of = [](const Member &M) -> uint64_t { return member_cost(M.de - M.ds, M.exp_isize); };
    uint64_t total_cost = 0;
    for (auto &M : mem)
        if (M.status == GZGPU_OK && !M.stored) total_cost += cost_of(M);
    uint64_t target_cost = std::max<uint64_t>(
        MIN_PIECE_OUT, (uint64_t)((double)total_cost / (PIECES_PER_SLOT * SLOT_FRACTION * (double)slots)));
    if (c->max_pieces) target_cost = std::max<uint64_t>(target_cost, total_cost / c->max_pieces + 1);
    const uint64_t ppc = c->tune.pieces_per_chunk;
    auto spacing_of = [&](const Member &M) -> uint64_t {
        if (ppc) return std::max<uint64_t>(chunk_bits / ppc, min_spacing);
        const uint64_t npieces = (cost_of(M) + target_cost - 1) / target_cost;
        if (npieces <= 1) return ~0ull;
        return std::max<uint64_t>((M.de - M.ds) / npieces, min_spacing_of(M));
    };

    // ---- 2/3: find (one pass over every member to be split) + plan
    auto splits = [&](const Member &M) {
        const uint64_t spacing = spacing_of(M);
        return M.status == GZGPU_OK && !M.stored && M.split && spacing != ~0ull && M.de > M.ds + spacing;
    };
    // Output estimates per block (k_bf_estimate) pay off where a member is
    // cut into many pieces: only there can content change along it enough to
    // matter.  Members cut into fewer are planned by their average.
    std::vector<std::pair<uint64_t, uint64_t>> ranges;
    std::vector<uint8_t> final_ok;
    bool estimate = false;
    std::vector<uint8_t> split_m(mem.size());
    for (uint32_t m = 0; m < mem.size(); m++) split_m[m] = splits(mem[m]);
    const auto skip = stored_payloads(c, d_in, in_len, mem, split_m);
    for (uint32_t m = 0; m < mem.size(); m++)
        if (split_m[m]) {
            const Member &M = mem[m];
            const uint64_t tail0 = M.de - std::min<uint64_t>(M.de - M.ds, FINAL_SCAN_BITS);
            // the member's bits less its stored payloads, cut at tail0 (the
            // tail also reports final-block headers)
            auto add = [&](uint64_t a, uint64_t b) {
                if (a < tail0 && a < b) {
                    ranges.push_back({a, std::min(b, tail0)});
                    final_ok.push_back(0);
                }
                if (b > tail0 && a < b) {
                    ranges.push_back({std::max(a, tail0), b});
                    final_ok.push_back(1);
                }
            };
            uint64_t a = M.ds;
            for (auto &iv : skip[m]) {
                if (iv.first > a) add(a, iv.first);
                a = std::max(a, iv.second);
            }
            add(a, M.de);
            c->st.stored_runs_skipped += (uint32_t)skip[m].size();
            estimate = estimate || (M.de - M.ds) / spacing_of(M) >= ESTIMATE_MIN_PIECES;
        }
    const double th0 = debug_enabled() ? host_ms() : 0;
    const std::vector<gzgpu_boundary> all_hits = find_hits(c, d_in, in_len, ranges, estimate, &final_ok);
    const double th1 = debug_enabled() ? host_ms() : 0;
    c->st.candidates += (uint32_t)all_hits.size();
    auto hit_lt = [](const gzgpu_boundary &a, uint64_t bit) { return a.bit_offset < bit; };
    std::vector<PieceHost> P;
    P.reserve(mem.size() + (size_t)std::min<uint64_t>(total_cost / target_cost, 1u << 24) + 16);
    // per-member planning scratch, reused across members
    std::vector<uint64_t> sx, gpts;           // stretch starts (bits) then the member end; grid points
    std::vector<double> sent, scost, cc, ce;  // entries / cost per bit of each stretch; their running sums
    std::vector<uint8_t> sstored, shadow, chain_ok;
    for (uint32_t m = 0; m < mem.size(); m++) {
        Member &M = mem[m];
        if (M.status != GZGPU_OK || M.stored) continue;
        PieceHost p0;
        p0.d.start_bit = M.ds;
        p0.d.stop_bit = NO_STOP;
        p0.d.limit_bit = M.limit;
        p0.d.dst = M.out;
        p0.d.cap = M.out_cap;
        p0.d.window = nullptr;
        p0.d.flags = PF_DIRECT;
        p0.d.member = m;
        p0.member = m;
        p0.acc_lo = p0.acc_hi = M.ds;
        p0.confirmed = true;
        M.pieces.clear();
        M.pieces.push_back((uint32_t)P.size());
        P.push_back(p0);
        if (!splits(M)) continue;
        const double t_mem0 = debug_level() >= 2 ? host_ms() : 0;
        const uint64_t spacing = spacing_of(M);
        struct {
            const gzgpu_boundary *b, *e;
            size_t size() const { return (size_t)(e - b); }
            const gzgpu_boundary &operator[](size_t i) const { return b[i]; }
            const gzgpu_boundary *begin() const { return b; }
            const gzgpu_boundary *end() const { return e; }
        } hits{all_hits.data() + (std::lower_bound(all_hits.begin(), all_hits.end(), M.ds, hit_lt) - all_hits.begin()),
               all_hits.data() + (std::lower_bound(all_hits.begin(), all_hits.end(), M.de, hit_lt) - all_hits.begin())};
        // piece starts: per grid point a hit within an eighth of a spacing
        // of it, else a start inside the block that holds the grid point
        // (MIDBLOCK: the piece reads that block's header, at the last hit
        // before the point, and starts at the point's sync point), else the
        // hit nearest to it within half a spacing either way (pieces come out
        // near equal, and the last one to finish sets the launch time).
        // Preference among hits: a dynamic-block hit;
        // else a stored hit whose payload ends where another hit's accepted
        // range begins (a final stored block's range starts one bit late, past
        // its BFINAL bit) or at the member's end.  That rules out LEN/~NLEN
        // coincidences inside stored data and the copy of a full block's
        // LEN/NLEN seen one byte early (00 FF FF 00 00 also reads as LEN
        // 0xFF00, NLEN 0x00FF), which both cost a re-decode round when
        // chosen.  A scheduling choice only: any start is verified by the
        // predecessor's end.
        const uint64_t span = M.de - M.ds;
        const uint64_t np = std::max<uint64_t>(2, (span + spacing / 2) / spacing);
        size_t h = 0, hk = 0;  // hk: the first hit past the current grid point
        uint64_t last = M.ds;
        struct Start {
            uint64_t start, acc_lo, acc_hi, hdr;
            bool mid;
        };
        std::vector<Start> chosen;
        // The first hit at or after `bit`, searched forward from `from` (the
        // previous answer of the same pass: the targets mostly increase) in
        // doubling steps, then by bisection inside the last step; a target
        // before hits[from] takes a full search.  (A full binary search per
        // stored hit cost ~1 ms per member with 8k stored blocks among 20k
        // hits: cache misses over the whole hit array each time.)
        auto hit_seek = [&](size_t from, uint64_t bit) -> size_t {
            const size_t N = hits.size();
            from = std::min(from, N);
            if (from > 0 && hits[from - 1].bit_offset >= bit)
                return (size_t)(std::lower_bound(hits.begin(), hits.end(), bit, hit_lt) - hits.begin());
            size_t lo = from, hi = from, step = 1;
            while (hi < N && hits[hi].bit_offset < bit) {
                lo = hi + 1;
                hi += step;
                step <<= 1;
            }
            hi = std::min(hi, N);
            return (size_t)(std::lower_bound(hits.begin() + lo, hits.begin() + hi, bit, hit_lt) - hits.begin());
        };
        // chained[i]: stored hit i's payload ends where another hit's
        // accepted range begins, or at the member's end
        chain_ok.assign(hits.size(), 0);
        size_t cur_chain = 0;
        for (size_t i = 0; i < hits.size(); i++) {
            if (hits[i].kind != 0) continue;
            const uint64_t o = hits[i].seek_bit + 3;  // the LEN field
            const uint64_t next = (o / 8 + 4 + (hits[i].info >> 16)) * 8;
            bool ok = next == M.de;
            cur_chain = hit_seek(cur_chain, next > 16 ? next - 16 : 0);
            for (size_t k = cur_chain; !ok && k < hits.size() && hits[k].bit_offset <= next + 1; ++k)
                ok = hits[k].seek_bit >= next;
            chain_ok[i] = ok;
        }
        auto chained = [&](size_t i) { return chain_ok[i] != 0; };
        // Dynamic-header hits inside the payload of a chained stored block
        // are data, not blocks of this stream (a gzip of deflate data: .gz,
        // .zip, .png): every piece started there would miss.  They are
        // shadowed for planning.
        shadow.assign(hits.size(), 0);
        size_t cur_pay = 0;
        for (size_t i = 0; i < hits.size(); i++) {
            if (hits[i].kind != 0 || !chained(i)) continue;
            const uint64_t o = hits[i].seek_bit + 3;
            const uint64_t pay = (o / 8 + 4) * 8, next = (o / 8 + 4 + (hits[i].info >> 16)) * 8;
            cur_pay = hit_seek(cur_pay, pay);
            for (size_t k = cur_pay; k < hits.size() && hits[k].bit_offset < next; ++k)
                if (hits[k].kind == 2) shadow[k] = 1;
        }
        auto dyn_hit = [&](size_t j) { return hits[j].kind == 2 && !shadow[j]; };
        // a BFINAL = 1 header: a block to start inside of, never a piece start
        // (the stop rule stops only at non-final headers)
        auto final_hit = [&](size_t j) { return hits[j].kind == 2 && ((hits[j].info >> HIT_FINAL_SHIFT) & 1u); };
        // Decode cost and output along the member, per stretch between
        // consecutive real hits (dynamic headers, chained stored blocks):
        // a stored block's output is its LEN; a dynamic block's is estimated
        // from its code lengths (entries per bit, from the block finder) and
        // the estimates are scaled so that the member's total matches its
        // trailer ISIZE.  Cost per bit as in member_cost: 5 + entries per bit
        // for Huffman blocks, 1/16 for stored ones.  Mixed content (stored
        // runs between text and u32 stretches) then gets pieces of equal
        // decode time, not of equal compressed size.  On the 2 GiB mixed
        // member (against max(4, entries per bit) and 1/32) the inflate
        // launch went 17.2 -> 14.9 ms at L1 and 14.5 -> 12.2 at L9; a
        // constant of 4 or 6, or stored weights of 1/32, 1/12 or 1/8, were
        // slower on one level or both.
        // (the first stretch, from the member start to its first real hit
        // or at it, takes that hit's kind: a member's first block is almost
        // always of the kind of its second)
        sx.clear();
        sent.clear();
        scost.clear();
        sstored.clear();
        {
            const double avg = (double)M.exp_isize / (double)std::max<uint64_t>(1, M.de - M.ds);
            auto ratio_of = [&](size_t j) { return estimate ? ratio_table()[hits[j].info >> 20] : avg; };
            for (size_t j = 0; j < hits.size(); j++) {
                const bool st_hit = hits[j].kind == 0 && chained(j);
                if (!dyn_hit(j) && !st_hit) continue;
                const double v = st_hit ? (double)(hits[j].info >> 16) : ratio_of(j);  // stored: LEN, per bit below
                if (sx.empty()) {
                    sx.push_back(M.ds);
                    sstored.push_back(st_hit ? 1 : 0);
                    // the first block's LEN is unknown: a stored first stretch is
                    // costed as bytes (8 bits per entry) until the ISIZE fit
                    sent.push_back(st_hit ? -1.0 : v);
                    if (hits[j].bit_offset <= M.ds) {
                        sent.back() = v;
                        continue;
                    }
                }
                if (hits[j].bit_offset <= sx.back()) continue;
                sx.push_back(hits[j].bit_offset);
                sstored.push_back(st_hit ? 1 : 0);
                sent.push_back(v);
            }
            if (sx.empty()) {
                sx.push_back(M.ds);
                sstored.push_back(0);
                sent.push_back(avg);
            }
            sx.push_back(M.de);
            for (size_t j = 0; j + 1 < sx.size(); j++)
                if (sent[j] < 0) sent[j] = (double)(sx[j + 1] - sx[j]) / 8.0;
            double stored_ent = 0, dyn_est = 0;
            for (size_t j = 0; j + 1 < sx.size(); j++) {
                const double bits = (double)(sx[j + 1] - sx[j]);
                if (sstored[j]) {
                    stored_ent += sent[j];
                    sent[j] = bits > 0 ? sent[j] / bits : 0;
                } else {
                    dyn_est += sent[j] * bits;
                }
            }
            const double want = (double)M.exp_isize - stored_ent;
            const double scale = dyn_est > 0 && want > 0 ? std::min(8.0, std::max(0.125, want / dyn_est)) : 1.0;
            scost.resize(sent.size());
            for (size_t j = 0; j + 1 < sx.size(); j++) {
                if (!sstored[j]) sent[j] *= scale;
                scost[j] = sstored[j] ? 1.0 / 16.0 : 5.0 + sent[j];
            }
        }
        const double t_str = debug_level() >= 2 ? host_ms() : 0;
        const size_t nseg = sx.size() - 1;
        cc.assign(nseg + 1, 0.0);  // cumulative cost / entries at stretch starts
        ce.assign(nseg + 1, 0.0);
        for (size_t j = 0; j < nseg; j++) {
            cc[j + 1] = cc[j] + scost[j] * (double)(sx[j + 1] - sx[j]);
            ce[j + 1] = ce[j] + sent[j] * (double)(sx[j + 1] - sx[j]);
        }
        // (cost, entries) at bit x; queries in increasing x per cursor
        struct Cursor {
            size_t j = 0;
        };
        auto cum_at = [&](Cursor &cur, uint64_t x, double &cost, double &ent) {
            x = std::min(std::max(x, M.ds), M.de);
            while (cur.j + 1 < nseg && sx[cur.j + 1] <= x) cur.j++;
            const double dx = (double)(x - sx[cur.j]);
            cost = cc[cur.j] + scost[cur.j] * dx;
            ent = ce[cur.j] + sent[cur.j] * dx;
        };
        // grid points: the starts of pieces 1 .. np-1.  Piece sizes fall
        // linearly along the member (density 1 - TILT x over x in [0, 1]):
        // pieces are pulled largest first, so the member's first pieces go
        // to the first-dispatched wavefronts, and those run fastest (the
        // SIMD issue arbiter favours its oldest wave: on a 1 GiB text member
        // the first fifth dispatched decoded at 86.5 MB/s, the last fifth at
        // 70).  (Cutting each piece into a large head and a small tail,
        // tails pulled last, measured slower: a tail piece is all marker
        // phase, and the window chain and marker resolution grow with the
        // piece count.)  Measured on that member: inflate kernel 4.33 ms
        // flat, 4.20 ms at a slope of 0.22-0.3.  With the issue-priority
        // balance inside the launch (progress_report in k_inflate.h) the
        // lagging wavefronts issue first, and equal pieces are best (kernel
        // 3.76 ms sloped without it, 3.61 ms flat with it).
        const double tilt = prog_balance() ? 0.0 : TILT;
        gpts.resize(np + 1);  // grid point k: where the cumulative cost reaches its share
        {
            size_t j = 0;
            for (uint64_t k = 0; k <= np; k++) {
                const double x = (double)k / (double)np;
                const double C = cc[nseg] * (x - 0.5 * tilt * x * x) / (1.0 - 0.5 * tilt);
                while (j + 1 < nseg && cc[j + 1] <= C) j++;
                const double b = (double)sx[j] + (C - cc[j]) / std::max(scost[j], 1e-12);
                gpts[k] = std::min<uint64_t>(M.de, std::max<uint64_t>(M.ds, (uint64_t)b));
            }
        }
        auto grid_at = [&](uint64_t k) { return gpts[k]; };
        for (uint64_t gi = 1; gi < np; gi++) {
            const uint64_t g = grid_at(gi);
            if (g <= last) continue;
            const uint64_t rest = std::min(g - grid_at(gi - 1), grid_at(gi + 1) - g);
            const uint64_t lo = std::max<uint64_t>(last + 1, g - rest / 2), hi = g + rest / 2;
            while (h < hits.size() && hits[h].bit_offset < lo) h++;
            size_t best = hits.size();
            uint64_t bestd = ~0ull;
            for (int pass = 0; pass < 2 && best == hits.size(); pass++)
                for (size_t j = h; j < hits.size() && hits[j].bit_offset < hi; j++) {
                    const bool ok = pass == 0 ? dyn_hit(j) && !final_hit(j) : hits[j].kind == 0 && chained(j);
                    if (!ok) continue;
                    const uint64_t d = hits[j].bit_offset > g ? hits[j].bit_offset - g : g - hits[j].bit_offset;
                    if (d < bestd) {
                        bestd = d;
                        best = j;
                    }
                }
            if (midblock_enabled() && !c->tune.miss_every && (best == hits.size() || bestd > rest / 8)) {
                // the block holding g starts at the last hit at or before g
                // (the member's first block when there is none); a start
                // needs MID_HEAD bits after that header hit, for the header
                // itself, and MID_TAIL bits before the next hit.  Stored hits
                // that are not chained are LEN/NLEN coincidences and skipped.
                auto real = [&](size_t j) { return dyn_hit(j) || (hits[j].kind == 0 && chained(j)); };
                while (hk < hits.size() && hits[hk].bit_offset <= g) hk++;  // grid points increase
                size_t k = hk;
                size_t prev = k;
                while (prev > 0 && !real(prev - 1)) prev--;
                while (k < hits.size() && !real(k)) k++;
                const uint64_t hdr = prev ? hits[prev - 1].seek_bit : M.ds;
                const bool dyn = prev == 0 || hits[prev - 1].kind == 2;
                const uint64_t next = k < hits.size() ? hits[k].bit_offset : M.de;
                if (dyn && g >= hdr + MID_HEAD && g >= last + MID_HEAD && next >= g + MID_TAIL) {
                    if (prev && final_hit(prev - 1)) c->st.final_block_pieces++;
                    chosen.push_back({g, g, g, hdr, true});
                    last = g;
                    continue;
                }
            }
            if (best == hits.size()) continue;  // no boundary near this grid point
            if (hits[best].seek_bit >= M.de) break;
            chosen.push_back({hits[best].seek_bit, hits[best].bit_offset, hits[best].seek_bit, 0, false});
            last = hits[best].bit_offset;
            h = best + 1;
        }
        if (debug_level() >= 2)
            fprintf(stderr,
                    "[gzgpu] plan member %u: bits %llu isize %u cost %llu (by stretch %.0f) target %llu spacing %llu "
                    "grid %llu hits %zu stretches %zu chosen %zu; stretches %.3f ms, grid %.3f ms\n",
                    m, (unsigned long long)span, M.exp_isize, (unsigned long long)cost_of(M), cc[nseg],
                    (unsigned long long)target_cost, (unsigned long long)spacing, (unsigned long long)np, hits.size(),
                    nseg, chosen.size(), t_str - t_mem0, host_ms() - t_str);
        uint64_t tail_hdr = M.ds;  // header of the member's last real hit
        for (size_t j = hits.size(); j-- > 0;)
            if (dyn_hit(j) || (hits[j].kind == 0 && chained(j))) {
                tail_hdr = hits[j].seek_bit;
                break;
            }
        if (c->tune.miss_every)  // test hook: every n-th start one bit late, a certain miss
            for (size_t k = c->tune.miss_every - 1; k < chosen.size(); k += c->tune.miss_every)
                chosen[k].start = chosen[k].acc_lo = chosen[k].acc_hi = chosen[k].start + 1;
        Cursor cur_lo, cur_hi;
        for (size_t k = 0; k < chosen.size(); k++) {
            const Start &b = chosen[k];
            PieceDesc &prev = P[M.pieces.back()].d;
            prev.stop_bit = b.acc_lo;
            if (b.mid) {
                prev.flags |= PF_STOPMID;
                prev.stop_hdr = b.hdr;
            }
            PieceHost q;
            q.d.start_bit = b.start;
            q.d.stop_bit = NO_STOP;
            q.d.limit_bit = M.limit;
            q.d.window = nullptr;
            q.d.flags = PF_MARKER | (b.mid ? PF_MIDBLOCK : 0u);
            q.d.hdr_bit = b.mid ? b.hdr : 0;
            q.d.member = m;
            q.member = m;
            c->st.midblock_pieces += b.mid ? 1u : 0u;
            q.acc_lo = b.acc_lo;  // MIDBLOCK: the sync point, known after the launch
            q.acc_hi = b.acc_hi;
            // a start inside the block of the member's last real hit may lie
            // in the final block, which the finder never reports (it takes
            // BFINAL = 0 headers only): its sync then fails and the piece
            // before it runs to the member's end, so that piece's slab is
            // sized for the rest of the member (no overflow and no second
            // launch for it)
            const bool tail_next = k + 1 < chosen.size() && chosen[k + 1].mid && chosen[k + 1].hdr == tail_hdr;
            const uint64_t next = k + 1 < chosen.size() ? std::max(chosen[k + 1].acc_lo, b.acc_lo) : M.de;
            double c0, e0, c1, e1;
            cum_at(cur_lo, b.acc_lo, c0, e0);
            cum_at(cur_hi, next, c1, e1);
            const double est = (tail_next ? ce[nseg] : e1) - e0;
            q.d.cap = marker_slab_bytes((uint64_t)(1.4 * est) + 4096);
            q.plan_cost = (uint64_t)(256.0 * (c1 - c0)) + 1;
            if (k == 0) P[M.pieces.back()].plan_cost = (uint64_t)(256.0 * c0) + 1;
            M.pieces.push_back((uint32_t)P.size());
            P.push_back(q);
        }
    }
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    if (debug_enabled())
        fprintf(stderr, "[gzgpu] host: find %.3f ms (%zu hits), plan %.3f ms (%zu pieces)\n", th1 - th0, all_hits.size(),
                host_ms() - th1, P.size());

    const double tslab = debug_enabled() ? host_ms() : 0;
    // slab pool for the first launch
    uint64_t pool = 0;
    for (auto &p : P)
        if (p.d.flags & PF_MARKER) pool += p.d.cap;
    c->d_slab.ensure(std::max<uint64_t>(pool, 256));
    {
        uint64_t off = 0;
        for (auto &p : P)
            if (p.d.flags & PF_MARKER) {
                p.d.dst = c->d_slab.as<uint8_t>() + off;
                off += p.d.cap;
            }
    }

    if (debug_enabled()) fprintf(stderr, "[gzgpu] host: slabs %.3f ms\n", host_ms() - tslab);
    std::vector<PieceOut> R(P.size());
    std::vector<uint32_t> todo(P.size());
    for (uint32_t i = 0; i < P.size(); i++) todo[i] = i;
    c->st.pieces = (uint32_t)P.size();
    // linked[k]: piece k starts where the piece before it (the nearest one
    // not merged away) ended, by the latest results.  A member is decoded
    // when every piece is linked: by induction from piece 0 (the deflate
    // start) each piece then starts on the true chain of block boundaries.
    std::vector<uint8_t> linked(P.size(), 0), in_todo(P.size(), 0);
    int rounds = 0;
    double tpost = 0;
    while (!todo.empty()) {
        ++rounds;
        if (rounds > 1) c->st.redecoded += (uint32_t)todo.size();
        launch_inflate(c, d_in, in_len, P, mem, todo, R);
        if (debug_enabled()) tpost = host_ms();
        for (uint32_t pi : todo) in_todo[pi] = 0;
        todo.clear();
        auto push_todo = [&](uint32_t pi) {
            if (!in_todo[pi]) {
                in_todo[pi] = 1;
                todo.push_back(pi);
            }
        };

        // Speculation check (chunk_fetcher.rs:2506-2729 semantics: a start is
        // accepted only where the decode before it ends).  Every piece that
        // missed is re-decoded from its predecessor's end when that
        // predecessor stopped cleanly, all in the next launch, together with
        // the following pieces up to the next start that looks true: a run
        // of false starts costs one more launch, not one per piece.  A
        // re-decode from a wrong end is itself re-checked in the next round.
        // From round SPEC_FALLBACK_ROUND on, a miss takes the member's whole
        // remainder, so every valid stream finishes within a few launches.
        uint64_t extra_need = 0;
        std::vector<uint32_t> grow;
        for (uint32_t m = 0; m < mem.size(); m++) {
            Member &M = mem[m];
            if (M.status != GZGPU_OK || M.pieces.empty()) continue;
            const size_t np = M.pieces.size();
            // pass 1: local links against the current predecessor
            linked[M.pieces[0]] = 1;
            uint32_t prev = M.pieces[0];
            for (size_t k = 1; k < np; k++) {
                const uint32_t pi = M.pieces[k];
                PieceHost &p = P[pi];
                if (p.dropped) break;
                if (p.merged) continue;
                if (p.d.flags & PF_MIDBLOCK) {
                    // the start is the sync point the piece found; none: unlinkable
                    const PieceOut &me = R[pi];
                    const bool synced = !(me.status == PS_ERROR && me.err == E_SYNC);
                    p.acc_lo = synced ? me.start_bit : 1;
                    p.acc_hi = synced ? me.start_bit : 0;
                }
                const PieceOut &pr = R[prev];
                const bool now = pr.status == PS_STOPPED && pr.end_bit >= p.acc_lo && pr.end_bit <= p.acc_hi;
                if (now && !linked[pi] && (p.d.flags & PF_MIDBLOCK)) c->st.midblock_linked++;
                linked[pi] = now ? 1 : 0;
                prev = pi;
            }
            // pass 2: the end of the member, misses
            prev = M.pieces[0];
            bool chain = true;  // every piece so far linked
            for (size_t k = 1; k < np; k++) {
                const uint32_t pi = M.pieces[k];
                PieceHost &p = P[pi];
                if (p.dropped) break;
                if (p.merged) continue;
                const PieceOut &pr = R[prev];
                if (chain && pr.status == PS_FINAL) {
                    // the member ends in the piece before: nothing after it
                    for (size_t j = k; j < np; j++) P[M.pieces[j]].dropped = true;
                    break;
                }
                if (!linked[pi] && pr.status == PS_STOPPED && !in_todo[prev]) {
                    const uint64_t e = pr.end_bit;
                    // the group [k, j): re-decoded as one piece from e, up to
                    // the planned start of the next piece whose start looks
                    // true: its predecessor's decode ends there, or its own
                    // decode ran cleanly to its stop (a false start almost
                    // always meets an invalid code or distance first)
                    auto plausible = [&](size_t q) {
                        const PieceHost &h = P[M.pieces[q]];
                        const PieceOut &o = R[M.pieces[q]];
                        return !h.merged && h.acc_lo > e &&
                               (linked[M.pieces[q]] || o.status == PS_STOPPED || o.status == PS_FINAL);
                    };
                    size_t j = k + 1;
                    if (rounds < SPEC_FALLBACK_ROUND) {
                        while (j < np && !P[M.pieces[j]].dropped && !plausible(j)) j++;
                    } else {
                        j = np;
                    }
                    const bool to_end = j >= np || P[M.pieces[j]].dropped;
                    if (debug_enabled())
                        fprintf(stderr, "[gzgpu] miss: member %u piece %zu start [%llu,%llu] prev end %llu, redo %zu pieces\n",
                                m, k, (unsigned long long)p.acc_lo, (unsigned long long)p.acc_hi,
                                (unsigned long long)e, j - k);
                    const PieceHost &last = P[M.pieces[j - 1]];
                    uint64_t cap = 0;
                    for (size_t q = k; q < j; q++) cap += P[M.pieces[q]].d.cap;
                    p.d.stop_bit = to_end ? NO_STOP : last.d.stop_bit;
                    p.d.stop_hdr = last.d.stop_hdr;
                    p.d.flags = (p.d.flags & ~(PF_MIDBLOCK | PF_STOPMID)) | (to_end ? 0u : (last.d.flags & PF_STOPMID));
                    p.d.start_bit = p.acc_lo = p.acc_hi = e;
                    p.d.hdr_bit = 0;
                    p.plan_cost = 0;
                    for (size_t q = k + 1; q < j; q++) {
                        P[M.pieces[q]].merged = true;
                        linked[M.pieces[q]] = 1;
                    }
                    if (!to_end) linked[M.pieces[j]] = 0;  // re-checked against the new end
                    if (j > k + 1 || p.d.stop_bit == NO_STOP) {
                        // the group's range in one slab (the tail piece's slab
                        // already covers the member's rest)
                        const uint64_t want = to_end ? marker_slab_bytes((uint64_t)(1.3 * (double)(M.de - e) /
                                                                                     (double)std::max<uint64_t>(1, M.de - M.ds) *
                                                                                     (double)M.exp_isize))
                                                      : cap;
                        p.d.cap = std::max(want, p.d.cap);
                        extra_need += p.d.cap;
                        grow.push_back(pi);
                    }
                    linked[pi] = 1;  // starts at e by construction
                    push_todo(pi);
                    k = j - 1;
                    prev = pi;
                    chain = false;
                    continue;
                }
                chain = chain && linked[pi];
                prev = pi;
            }
            // errors and overflows: fatal inside the linked prefix; overflowed
            // speculative pieces are regrown (most are valid)
            bool prefix = true;
            for (size_t k = 0; k < np; k++) {
                const uint32_t pi = M.pieces[k];
                PieceHost &p = P[pi];
                if (p.dropped) break;
                if (p.merged) continue;
                prefix = prefix && linked[pi] && !in_todo[pi];
                if (in_todo[pi]) continue;
                const PieceOut &r = R[pi];
                if (prefix && r.status == PS_ERROR) {
                    M.status = GZGPU_DECODE_FAILED;
                    M.msg = "deflate stream invalid at bit " + std::to_string(r.end_bit) + " (piece " +
                            std::to_string(k) + ", error " + std::to_string(r.err) + ")";
                    break;
                }
                if (r.status == PS_OVERFLOW) {
                    if (p.d.flags & PF_DIRECT) {
                        M.status = GZGPU_OUTPUT_TOO_SMALL;
                        M.msg = "output capacity exceeded";
                        break;
                    }
                    uint64_t entries = r.n16 + r.n8;
                    uint64_t span = (p.d.stop_bit == NO_STOP ? M.de : p.d.stop_bit) - p.d.start_bit;
                    double rate = r.bits_done ? (double)entries / (double)r.bits_done : 1.0;
                    uint64_t want = (uint64_t)(rate * (double)span * 1.5) + entries;
                    p.d.cap = std::max<uint64_t>(marker_slab_bytes(want), 2 * p.d.cap);
                    extra_need += p.d.cap;
                    grow.push_back(pi);
                    push_todo(pi);
                }
            }
            if (M.status == GZGPU_OK && rounds >= SPEC_MAX_ROUNDS) {
                bool pending = false;
                for (size_t k = 0; k < np && !pending; k++) pending = in_todo[M.pieces[k]] != 0;
                if (pending) {
                    M.status = GZGPU_DECODE_FAILED;
                    M.msg = "speculation did not converge in " + std::to_string(rounds) + " launches";
                }
            }
            if (M.status != GZGPU_OK) {
                todo.erase(std::remove_if(todo.begin(), todo.end(),
                                          [&](uint32_t pi) {
                                              if (P[pi].member != m) return false;
                                              in_todo[pi] = 0;
                                              return true;
                                          }),
                           todo.end());
                grow.erase(std::remove_if(grow.begin(), grow.end(), [&](uint32_t pi) { return P[pi].member == m; }),
                           grow.end());
            }
        }
        if (!grow.empty()) {
            extra_need = 0;
            for (uint32_t pi : grow) extra_need += P[pi].d.cap;
            DevBuf b;
            b.ensure(extra_need);
            uint64_t off = 0;
            for (uint32_t pi : grow) {
                P[pi].d.dst = b.as<uint8_t>() + off;
                off += P[pi].d.cap;
            }
            c->extra_slabs.push_back(b);
        }
    }
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    const double tlink = debug_enabled() ? host_ms() : 0;

    // ---- assemble: offsets, sizes, resolve lists
    std::vector<ResolvePiece> rps;
    std::vector<uint32_t> chains;
    std::vector<uint32_t> tile_first;  // per marker piece: its first body tile (k_resolve_body)
    uint32_t ntiles = 0;
    std::vector<uint32_t> tile_lo(mem.size(), 0), tile_hi(mem.size(), 0);  // each member's tiles
    std::vector<uint32_t> rp_lo(mem.size(), 0), rp_hi(mem.size(), 0);
    struct Out {
        uint32_t pi;
        uint64_t pos, L;
    };
    std::vector<Out> outs;  // one member's pieces with output, in order
    for (uint32_t m = 0; m < mem.size(); m++) {
        Member &M = mem[m];
        if (M.status != GZGPU_OK) continue;
        tile_lo[m] = tile_hi[m] = ntiles;
        rp_lo[m] = rp_hi[m] = (uint32_t)rps.size();
        if (M.stored) continue;  // copied whole by k_stored_copy; out_len is known
        // pass 1: sizes, offsets, status
        outs.clear();
        uint64_t pos = 0;
        bool final_seen = false;
        for (size_t k = 0; k < M.pieces.size(); k++) {
            uint32_t pi = M.pieces[k];
            if (P[pi].dropped) break;
            if (P[pi].merged) continue;
            if (!linked[pi]) {
                M.status = GZGPU_DECODE_FAILED;
                M.msg = "speculative piece " + std::to_string(k) + " never linked to its predecessor";
                break;
            }
            const PieceOut &r = R[pi];
            const uint64_t L = r.n16 + r.n8;
            if ((P[pi].d.flags & PF_MARKER) && pos + L > M.out_cap) {
                M.status = GZGPU_OUTPUT_TOO_SMALL;
                M.msg = "output capacity exceeded";
                break;
            }
            outs.push_back({pi, pos, L});
            if (c->cp_collect && r.cp_bit != NO_CP && r.cp_bit >= M.ds)
                c->cp_collect->push_back({r.cp_bit - M.ds, pos + r.cp_out});
            pos += L;
            if (r.status == PS_FINAL) {
                final_seen = true;
                M.end_bit = r.end_bit;
                break;
            }
        }
        if (M.status != GZGPU_OK) continue;
        if (!final_seen) {
            M.status = GZGPU_DECODE_FAILED;
            M.msg = "deflate stream ended without a final block";
            continue;
        }
        M.out_len = pos;
        // pass 2: resolve lists.  A marker piece that flipped to u8 has final
        // bytes as its last 32768 entries.  When all of those are u8 entries
        // and the next piece with output holds 32768 entries or more (so no
        // later window reaches back into it), k_resolve_body resolves it
        // whole and the pieces up to that next one read their window from
        // its slab.  Every other marker piece is in a chain: k_chain writes
        // its last 32 KiB before k_resolve_body runs, so every window read
        // from the output is final by then.  A chain starts anew at a piece
        // that flipped (its end window needs no predecessor); chains run in
        // parallel, one workgroup each.
        bool chain_open = false;
        const uint8_t *flip_end = nullptr;  // the slab-resolved piece windows come from
        for (size_t k = 0; k < outs.size(); k++) {
            const uint32_t pi = outs[k].pi;
            const PieceOut &r = R[pi];
            const uint64_t L = outs[k].L;
            if (!(P[pi].d.flags & PF_MARKER)) {
                if (L) flip_end = nullptr;
                continue;
            }
            uint64_t nextL = WIN;  // the next piece with output (none: nothing reads back)
            for (size_t j = k + 1; j < outs.size(); j++)
                if (outs[j].L) {
                    nextL = outs[j].L;
                    break;
                }
            const bool self_windowed = r.u8_off != 0 && L >= WIN;
            const bool whole = self_windowed && r.n8 >= WIN && nextL >= WIN;
            if (whole) {
                if (chain_open) chains.push_back((uint32_t)rps.size());
                chain_open = false;
            } else if (!chain_open || self_windowed) {
                if (chain_open) chains.push_back((uint32_t)rps.size());
                chains.push_back((uint32_t)rps.size());
                chain_open = true;
            }
            ResolvePiece rp;
            rp.slab = P[pi].d.dst;
            rp.mout = M.out;
            rp.out_pos = outs[k].pos;
            rp.n16 = r.n16;
            rp.n8 = r.n8;
            rp.u8_off = r.u8_off;
            rp.tail = M.out_len - (outs[k].pos + L);
            // every entry is the body pass's, the end window too: k_chain
            // writes it first for the successor's window, and the body pass
            // writes the same bytes again with their CRC folded in (a CRC pass
            // of its own over the windows ran at ~0.2 TB/s: 0.57 ms of a
            // 256 MiB member's 3.1 ms)
            rp.body = L;
            rp.win_end = flip_end ? flip_end : M.out + outs[k].pos;
            rp.win_out = flip_end ? 0u : 1u;
            rp.pad = 0;
            rps.push_back(rp);
            tile_first.push_back(ntiles);
            ntiles += (uint32_t)((rp.body + RESOLVE_TILE - 1) / RESOLVE_TILE);
            if (L) flip_end = whole ? P[pi].d.dst + r.u8_off + r.n8 : nullptr;
        }
        if (chain_open) chains.push_back((uint32_t)rps.size());
        tile_hi[m] = ntiles;
        rp_hi[m] = (uint32_t)rps.size();
    }
    tile_first.push_back(ntiles);
    if (debug_level() >= 3) {
        for (size_t i = 0; i < rps.size(); i++)
            fprintf(stderr, "[gzgpu] rp %zu out %llu n16 %llu n8 %llu u8_off %llu body %llu win_out %u\n", i,
                    (unsigned long long)rps[i].out_pos, (unsigned long long)rps[i].n16, (unsigned long long)rps[i].n8,
                    (unsigned long long)rps[i].u8_off, (unsigned long long)rps[i].body, rps[i].win_out);
        for (size_t i = 0; i + 1 < chains.size(); i += 2)
            fprintf(stderr, "[gzgpu] chain %u..%u\n", chains[i], chains[i + 1]);
    }

    const double tasm = debug_enabled() ? host_ms() : 0;
    // ---- CRC32 lists (every member's output size is known now)
    // k_resolve_body returns the CRC of every tile it writes; the segments
    // here cover the rest of each member (pieces decoded straight to the
    // output, and each marker piece's last 32 KiB, written by k_chain)
    std::vector<uint64_t> seg_off, seg_tail;
    std::vector<uint32_t> seg_len, mseg, mtile;
    std::vector<uint64_t> mlen;
    std::vector<uint32_t> crc_members;
    const uint8_t *base = nullptr;
    for (auto &M : mem)
        if (M.status == GZGPU_OK && (!base || M.out < base)) base = M.out;
    for (uint32_t m = 0; m < mem.size(); m++) {
        Member &M = mem[m];
        if (M.status != GZGPU_OK) continue;
        crc_members.push_back(m);
        mseg.push_back((uint32_t)seg_off.size());
        mtile.push_back(tile_lo[m]);
        mtile.push_back(tile_hi[m]);
        mlen.push_back(M.out_len);
        // output bytes [a, b) as segments with boundaries on 16-byte
        // addresses: every segment but a range's first and last is a full,
        // aligned 64 KiB frame
        auto add_range = [&](uint64_t a, uint64_t b) {
            const uint64_t lead = (uint64_t)((uintptr_t)(M.out + a) & 15);
            for (uint64_t o = a; o < b;) {
                const uint64_t l = std::min<uint64_t>(o == a ? CRC_SEG - lead : CRC_SEG, b - o);
                seg_off.push_back((uint64_t)(M.out - base) + o);
                seg_len.push_back((uint32_t)l);
                seg_tail.push_back(M.out_len - o - l);
                o += l;
            }
        };
        // the gaps between the member's tiles: a marker piece's tiles cover
        // its first `body` entries
        uint64_t at = 0;
        for (uint32_t i = rp_lo[m]; i < rp_hi[m]; i++) {
            if (rps[i].body == 0) continue;
            add_range(at, rps[i].out_pos);
            at = rps[i].out_pos + rps[i].body;
        }
        add_range(at, M.out_len);
    }
    if (!crc_members.empty()) mseg.push_back((uint32_t)seg_off.size());
    const uint32_t nseg = (uint32_t)seg_off.size(), nm = (uint32_t)crc_members.size();

    // long chains: slots of the parallel-prefix window chain
    std::vector<uint32_t> slot_rp, slot_pos;
    uint32_t max_chain = 0;
    for (size_t i = 0; i + 1 < chains.size(); i += 2) max_chain = std::max(max_chain, chains[i + 1] - chains[i]);
    const bool chain_scan = max_chain > chain_scan_min();
    if (chain_scan) {
        for (size_t i = 0; i + 1 < chains.size(); i += 2)
            for (uint32_t q = chains[i]; q < chains[i + 1]; q++) {
                slot_rp.push_back(q);
                slot_pos.push_back(q - chains[i]);
            }
    }

    // ---- every resolve and CRC list in one staging buffer and one copy:
    // small copies between the kernels each left the GPU idle for a while
    struct Part {
        const void *src;
        size_t bytes, off;
    };
    Part parts[] = {{rps.data(), rps.size() * sizeof(ResolvePiece), 0}, {chains.data(), chains.size() * 4, 0},
                    {tile_first.data(), tile_first.size() * 4, 0}, {seg_off.data(), seg_off.size() * 8, 0},
                    {seg_len.data(), seg_len.size() * 4, 0}, {seg_tail.data(), seg_tail.size() * 8, 0},
                    {mseg.data(), mseg.size() * 4, 0}, {mlen.data(), mlen.size() * 8, 0},
                    {mtile.data(), mtile.size() * 4, 0},     {slot_rp.data(), slot_rp.size() * 4, 0},
                    {slot_pos.data(), slot_pos.size() * 4, 0}};
    size_t lists_bytes = 0;
    for (auto &pt : parts) {
        pt.off = lists_bytes;
        lists_bytes += (pt.bytes + 255) & ~(size_t)255;
    }
    if (lists_bytes) {
        c->d_lists.ensure(lists_bytes);
        uint8_t *h = (uint8_t *)c->pin.alloc(lists_bytes);
        for (auto &pt : parts)
            if (pt.bytes) memcpy(h + pt.off, pt.src, pt.bytes);
        HIPCHK(hipMemcpyAsync(c->d_lists.p, h, lists_bytes, hipMemcpyHostToDevice, c->stream));
    }
    auto part = [&](int i) { return (void *)(c->d_lists.as<uint8_t>() + parts[i].off); };
    if (debug_enabled())
        fprintf(stderr, "[gzgpu] host: link %.3f ms, assemble %.3f ms, crc lists + stage %.3f ms (%zu KB)\n",
                tlink - tpost, tasm - tlink, host_ms() - tasm, lists_bytes >> 10);
    const bool streaming = c->stream_sink && mem.size() == 1 && mem[0].status == GZGPU_OK && !mem[0].stored &&
                           ntiles > 0 && mem[0].out_len >= STREAM_MIN;
    std::vector<std::pair<uint64_t, int>> ready;  // (final output prefix, event) per resolve range
    if (!rps.empty()) {
        const uint32_t nch = (uint32_t)(chains.size() / 2);  // (first, end) pairs
        HIPCHK(hipEventRecord(c->ev[7], c->stream));
        if (nch && chain_scan) {
            const uint32_t ns = (uint32_t)slot_rp.size();
            c->d_scan.ensure((size_t)ns * WIN * 2 * 2);
            uint16_t *wb[2] = {c->d_scan.as<uint16_t>(), c->d_scan.as<uint16_t>() + (size_t)ns * WIN};
            const dim3 g(std::min<uint32_t>(ns, (uint32_t)c->num_cus * 8));
            c->st.launches += launch_chain_scan(c->stream, g, (const ResolvePiece *)part(0), (const uint32_t *)part(9),
                                                (const uint32_t *)part(10), ns, max_chain, wb);
            HIPCHK(hipGetLastError());
        } else if (nch) {
            hipLaunchKernelGGL(k_chain, dim3(std::min<uint32_t>(nch, (uint32_t)c->num_cus * 4)), dim3(1024), 0,
                               c->stream, (const ResolvePiece *)part(0), (const uint32_t *)part(1), nch);
            HIPCHK(hipGetLastError());
            c->st.launches++;
        }
        HIPCHK(hipEventRecord(c->ev[8], c->stream));
        if (ntiles) {
            ensure_crc_tables(c);
            c->d_tile_crc.ensure((size_t)ntiles * 4);
            // streaming: the tiles in ranges, in output order; after range r
            // every output byte before the first tile of range r + 1 is final
            // (the rest of the member is written by the inflate and k_chain)
            const uint32_t nr = streaming ? std::min<uint32_t>(STREAM_RANGES, ntiles) : 1;
            for (uint32_t r = 0; r < nr; r++) {
                const uint32_t t0 = (uint32_t)((uint64_t)ntiles * r / nr), t1 = (uint32_t)((uint64_t)ntiles * (r + 1) / nr);
                hipLaunchKernelGGL(k_resolve_body, dim3((uint32_t)std::min<size_t>(t1 - t0, (size_t)c->num_cus * 32)),
                                   dim3(256), 0, c->stream, (const ResolvePiece *)part(0), (const uint32_t *)part(2),
                                   (uint32_t)rps.size(), c->d_crc_tab.as<uint32_t>(), c->d_tile_crc.as<uint32_t>(),
                                   t0, t1);
                HIPCHK(hipGetLastError());
                c->st.launches++;
                if (streaming) {
                    HIPCHK(hipEventRecord(c->res_ev[r], c->stream));
                    uint64_t prefix = mem[0].out_len;
                    if (r + 1 < nr) {
                        const size_t k = (size_t)(std::upper_bound(tile_first.begin(), tile_first.end() - 1, t1) -
                                                  tile_first.begin()) - 1;
                        prefix = rps[k].out_pos + (uint64_t)(t1 - tile_first[k]) * RESOLVE_TILE;
                    }
                    ready.push_back({prefix, (int)r});
                }
            }
        }
    }
    HIPCHK(hipEventRecord(c->ev[3], c->stream));

    // ---- CRC32 per member
    if (!crc_members.empty()) {
        ensure_crc_tables(c);
        c->d_seg_crc.ensure(std::max<size_t>(nseg, 1) * 4);
        if (nseg) {
            hipLaunchKernelGGL(k_crc_segments, dim3(std::min<uint32_t>(nseg, (uint32_t)c->num_cus * 8)), dim3(256), 0,
                               c->stream, base, (const uint64_t *)part(3), (const uint32_t *)part(4),
                               (const uint64_t *)part(5), nseg, c->d_crc_tab.as<uint32_t>(),
                               c->d_seg_crc.as<uint32_t>());
            HIPCHK(hipGetLastError());
            c->st.launches++;
        }
        c->d_mcrc.ensure(nm * 4);
        c->d_tile_crc.ensure(std::max<size_t>(ntiles, 1) * 4);
        hipLaunchKernelGGL(k_crc_members, dim3(std::min<uint32_t>(nm, (uint32_t)c->num_cus * 8)), dim3(256), 0, c->stream,
                           c->d_seg_crc.as<uint32_t>(), (const uint32_t *)part(6), (const uint64_t *)part(7), nm,
                           c->d_mcrc.as<uint32_t>(), (const uint32_t *)c->d_tile_crc.p, (const uint32_t *)part(8));
        HIPCHK(hipGetLastError());
        c->st.launches++;
        const uint32_t *crcs = download<uint32_t>(c, c->d_mcrc.p, nm);
        HIPCHK(hipEventRecord(c->ev[4], c->stream));
        if (streaming) {
            // the output reaches the sink range by range while later ranges
            // resolve and the CRC runs; a CRC or size error is reported after
            // (partial output before a terminal error, single_member.rs:17-22)
            c->stream_rc = drain_ranges(c, mem[0].out, mem[0].out_len, ready);
            c->streamed = true;
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        for (uint32_t i = 0; i < nm; i++) {
            Member &M = mem[crc_members[i]];
            M.crc = crcs[i];
            if ((uint32_t)M.out_len != M.exp_isize) {
                M.status = GZGPU_SIZE_MISMATCH;
                M.msg = "ISIZE " + std::to_string(M.exp_isize) + " != decoded " + std::to_string(M.out_len);
            } else if (M.crc != M.exp_crc) {
                M.status = GZGPU_CRC_MISMATCH;
                char b[96];
                snprintf(b, sizeof b, "CRC32 %08x != trailer %08x", M.crc, M.exp_crc);
                M.msg = b;
            }
        }
    } else {
        HIPCHK(hipEventRecord(c->ev[4], c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (debug_enabled() && !rps.empty())
        fprintf(stderr, "[gzgpu] resolve: chain %.3f ms (%zu pieces, longest chain %u%s), body %.3f ms (%zu tiles)\n",
                ev_ms(c->ev[7], c->ev[8]), rps.size(), max_chain, chain_scan ? ", parallel prefix" : "",
                ev_ms(c->ev[8], c->ev[3]), (size_t)ntiles);
    c->st.ms_find += ev_ms(c->ev[0], c->ev[1]);
    c->st.ms_inflate += ev_ms(c->ev[1], c->ev[2]);
    c->st.ms_resolve += ev_ms(c->ev[2], c->ev[3]);
    c->st.ms_crc += ev_ms(c->ev[3], c->ev[4]);
    c->st.ms_total += ev_ms(c->ev[0], c->ev[4]);
    for (auto &b : c->extra_slabs) b.release();
    c->extra_slabs.clear();
}

// Parse envelope from host-visible bytes of one member starting at `p`.
int member_from_host(const uint8_t *p, size_t avail, uint64_t abs_off, Member &M, std::string &msg) {
    size_t hl;
    int st = parse_header(p, avail, &hl);
    if (st) {
        msg = "invalid gzip header";
        return st;
    }
    if (avail < hl + 8) {
        msg = "input shorter than header + trailer";
        return GZGPU_INVALID_FORMAT;
    }
    M.ds = (abs_off + hl) * 8;
    M.de = (abs_off + avail - 8) * 8;
    M.limit = (abs_off + avail) * 8;
    M.exp_crc = le32(p + avail - 8);
    M.exp_isize = le32(p + avail - 4);
    return GZGPU_OK;
}

int finish(gzgpu_ctx *c, int st, const std::string &msg) {
    if (st != GZGPU_OK) set_err(c, msg);
    return st;
}

template <typename F>
int guarded(gzgpu_ctx *c, F &&f, bool reset_staging = true) {
    try {
        // every call starts with an idle stream, so staging memory is free
        if (c && c->stream) {
            HIPCHK(hipStreamSynchronize(c->stream));
            if (reset_staging) c->pin.reset();
        }
        return f();
    } catch (const HipError &e) {
        set_err(c, std::string("HIP error: ") + hipGetErrorString(e.e) + " in " + e.what);
        (void)hipGetLastError();
        return GZGPU_GPU_UNAVAILABLE;
    } catch (const StatusError &e) {
        set_err(c, e.msg);
        return e.status;
    } catch (const std::bad_alloc &) {
        set_err(c, "host allocation failed");
        return GZGPU_OUT_OF_DEVICE_MEMORY;
    }
}

void init_constants(gzgpu_ctx *c) {
    const CrcConst &cc = crc_const();
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(gz::c_xbyte), cc.xbyte, sizeof(cc.xbyte)));
    (void)c;
}

// Hand `n` bytes of device output to the sink in order, in DRAIN_CHUNK
// pieces through two pinned staging buffers: the D2H copy of the next piece
// runs while the sink takes the current one (in-order emission while the
// next chunk is in flight, drain_one_pending chunk_fetcher.rs:2847).
constexpr uint64_t DRAIN_CHUNK = 8 * MIB;
template <typename Sink>
int drain_to_sink(gzgpu_ctx *c, const uint8_t *d, uint64_t n, Sink &sink) {
    if (!n) return 0;
    for (int i = 0; i < 2; i++) {
        if (!c->drain[i]) {
            if (hipHostMalloc((void **)&c->drain[i], DRAIN_CHUNK, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                c->drain[i] = nullptr;
                throw StatusError{GZGPU_GPU_UNAVAILABLE, "hipHostMalloc(output staging) failed"};
            }
        }
        if (!c->drain_ev[i]) HIPCHK(hipEventCreateWithFlags(&c->drain_ev[i], hipEventDisableTiming));
    }
    const uint64_t chunks = (n + DRAIN_CHUNK - 1) / DRAIN_CHUNK;
    auto enqueue = [&](uint64_t i) {
        const uint64_t off = i * DRAIN_CHUNK, len = std::min(DRAIN_CHUNK, n - off);
        HIPCHK(hipMemcpyAsync(c->drain[i & 1], d + off, len, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipEventRecord(c->drain_ev[i & 1], c->stream));
    };
    enqueue(0);
    int rc = 0;
    for (uint64_t i = 0; i < chunks; i++) {
        // buffer (i + 1) & 1 held chunk i - 1, which the sink has finished
        if (i + 1 < chunks) enqueue(i + 1);
        HIPCHK(hipEventSynchronize(c->drain_ev[i & 1]));
        if (!rc) rc = sink(c->drain[i & 1], (size_t)std::min(DRAIN_CHUNK, n - i * DRAIN_CHUNK));
    }
    return rc;
}

// Hand the member's output to the context's streaming sink in order, in
// DRAIN_CHUNK pieces: the D2H of a piece waits (on the copy stream) for the
// resolve range that finishes it, so it overlaps the later ranges and the
// CRC; two pinned buffers let the sink take one piece while the next copies.
int drain_ranges(gzgpu_ctx *c, const uint8_t *d, uint64_t n, const std::vector<std::pair<uint64_t, int>> &ready) {
    if (!n) return 0;
    for (int i = 0; i < 2; i++) {
        if (!c->drain[i]) {
            if (hipHostMalloc((void **)&c->drain[i], DRAIN_CHUNK, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                c->drain[i] = nullptr;
                throw StatusError{GZGPU_GPU_UNAVAILABLE, "hipHostMalloc(output staging) failed"};
            }
        }
        if (!c->drain_ev[i]) HIPCHK(hipEventCreateWithFlags(&c->drain_ev[i], hipEventDisableTiming));
    }
    const uint64_t chunks = (n + DRAIN_CHUNK - 1) / DRAIN_CHUNK;
    size_t r = 0;  // the first range not yet waited for
    int waited = -1;
    auto enqueue = [&](uint64_t i) {
        const uint64_t off = i * DRAIN_CHUNK, len = std::min(DRAIN_CHUNK, n - off);
        while (r < ready.size() && ready[r].first < off + len) r++;
        const int ev = ready[std::min(r, ready.size() - 1)].second;
        if (ev > waited) {
            HIPCHK(hipStreamWaitEvent(c->copy_stream, c->res_ev[ev], 0));
            waited = ev;
        }
        HIPCHK(hipMemcpyAsync(c->drain[i & 1], d + off, len, hipMemcpyDeviceToHost, c->copy_stream));
        HIPCHK(hipEventRecord(c->drain_ev[i & 1], c->copy_stream));
    };
    enqueue(0);
    int rc = 0;
    for (uint64_t i = 0; i < chunks; i++) {
        if (i + 1 < chunks) enqueue(i + 1);
        HIPCHK(hipEventSynchronize(c->drain_ev[i & 1]));
        if (!rc) rc = (*c->stream_sink)(c->drain[i & 1], (size_t)std::min(DRAIN_CHUNK, n - i * DRAIN_CHUNK));
    }
    return rc;
}

// Compressed size of the member at `p`, from its FEXTRA subfield: gzippy's
// "GZ" (4 bytes: the whole member; 2 bytes: size - 1; gz_member_len,
// src/decompress/bgzf.rs:2322-2357) or BGZF's "BC" (2 bytes: size - 1).
// 0 when the member carries neither.
uint64_t sized_member_len(const uint8_t *p, size_t avail) {
    if (avail < 18 || p[0] != 0x1f || p[1] != 0x8b || p[2] != 8 || !(p[3] & 4)) return 0;
    const size_t xlen = (size_t)p[10] | (size_t)p[11] << 8;
    if (12 + xlen > avail) return 0;
    const uint8_t *x = p + 12;
    for (size_t pos = 0; pos + 4 <= xlen;) {
        const size_t sl = (size_t)x[pos + 2] | (size_t)x[pos + 3] << 8;
        const bool gz = x[pos] == 'G' && x[pos + 1] == 'Z', bc = x[pos] == 'B' && x[pos + 1] == 'C';
        if (gz && sl == 4 && pos + 8 <= xlen) return le32(x + pos + 4);
        if ((gz || bc) && sl == 2 && pos + 6 <= xlen) return ((uint64_t)x[pos + 4] | (uint64_t)x[pos + 5] << 8) + 1;
        if (gz) return 0;
        pos += 4 + sl;
    }
    return 0;
}

// Member starts (and the end) of an input whose members all carry their size
// and tile it exactly (gz_coverage_is_pure, src/decompress/bgzf.rs:2295);
// empty when any member does not.
std::vector<uint64_t> sized_member_walk(const uint8_t *in, size_t n) {
    std::vector<uint64_t> offs;
    for (size_t off = 0; off < n;) {
        const uint64_t len = sized_member_len(in + off, n - off);
        if (len < 18 || len > n - off) return {};
        offs.push_back(off);
        off += len;
    }
    offs.push_back(n);
    return offs;
}

// Member spans of a gzip stream: candidate starts at 1f 8b 08 with clear
// reserved flag bits, each span running to the next candidate (the
// reference's member-boundary scan, scan_member_boundaries_fast,
// src/decompress/bgzf.rs:2090).  A candidate inside a member's payload cuts
// it short; decoding finds that out (the member does not end at its span's
// trailer).
std::vector<std::pair<size_t, size_t>> member_spans(const uint8_t *p, size_t n) {
    std::vector<std::pair<size_t, size_t>> v;
    if (n < 18 || p[0] != 0x1f || p[1] != 0x8b) return v;
    size_t start = 0;
    for (size_t q = 18; q + 18 <= n; q++) {
        const uint8_t *m = (const uint8_t *)memchr(p + q, 0x1f, n - 18 - q + 1);
        if (!m) break;
        q = (size_t)(m - p);
        if (q + 18 > n) break;
        if (p[q + 1] == 0x8b && p[q + 2] == 8 && (p[q + 3] & 0xE0) == 0 && q - start >= 18) {
            v.push_back({start, q - start});
            start = q;
        }
    }
    v.push_back({start, n - start});
    return v;
}

// ISIZE of a member span, or ~0 when it cannot belong to that span: deflate
// expands at most 1032:1 (258-byte copies from 2-bit codes, the reference's
// MAX_DEFLATE_EXPANSION, chunk_data.rs:334), so a larger ISIZE means a false
// magic inside the payload cut the span short.
uint64_t span_isize(const uint8_t *p, size_t len) {
    if (len < 18) return ~0ull;
    const uint64_t isz = le32(p + len - 4);
    return isz > 1032ull * len + 64 ? ~0ull : isz;
}

// Decode the members of a host buffer (multi-member walk,
// src/decompress/mod.rs:725-741):
//   1. members that all carry their size (gzippy's parallel "GZ" format,
//      BGZF) decode together in one pipeline, like a batch
//      (decompress_bgzf_parallel, src/decompress/bgzf.rs:1874);
//   2. else, with more than one candidate member start, the candidate spans
//      decode together in one pipeline (MultiMemberPar, bgzf.rs:2370, on the
//      spans of scan_member_boundaries_fast, :2090), accepted when every
//      member ends exactly at its span's trailer;
//   3. else, or when any member of those failed, member by member: each
//      member first bounded by the next candidate start, widened to the rest
//      of the input only when it does not end there; this walk reports a
//      failure where it finds it.
template <typename Sink>
int decompress_host_stream(gzgpu_ctx *c, const uint8_t *in, size_t in_len, Sink &&sink, uint64_t *total,
                           gzgpu_stats *stats) {
    c->st = gzgpu_stats{};
    *total = 0;
    if (in_len < 2 || in[0] != 0x1f || in[1] != 0x8b) return GZGPU_OK;  // non-gzip: Ok(0)
    c->d_in.ensure(in_len + 64);
    HIPCHK(hipMemcpyAsync(c->d_in.p, in, in_len, hipMemcpyHostToDevice, c->stream));
    gzgpu_tuning tn;
    gzgpu_get_tuning(c, &tn);
    // members [start, end) decoded together; true (and the output drained)
    // when every one decoded, verified and, if `exact`, ended at its trailer
    auto together = [&](const std::vector<std::pair<size_t, size_t>> &sp, bool exact, const char *what,
                        int &sink_rc) -> bool {
        std::vector<Member> mem(sp.size());
        uint64_t out_total = 0;
        for (size_t i = 0; i < mem.size(); i++) {
            std::string msg;
            if (member_from_host(in + sp[i].first, sp[i].second, sp[i].first, mem[i], msg) != GZGPU_OK) return false;
            out_total += mem[i].exp_isize;
        }
        c->d_out.ensure(out_total + 64);
        uint64_t o = 0;
        for (size_t i = 0; i < mem.size(); i++) {
            Member &M = mem[i];
            M.out = c->d_out.as<uint8_t>() + o;
            M.out_cap = M.exp_isize;
            M.split = tn.batch_split_bytes == 0 || sp[i].second > tn.batch_split_bytes;
            o += M.exp_isize;
        }
        run_members(c, c->d_in.as<uint8_t>(), in_len, mem);
        size_t bad = 0;
        while (bad < mem.size() && mem[bad].status == GZGPU_OK &&
               (!exact || (mem[bad].end_bit + 7) / 8 == mem[bad].de / 8))
            bad++;
        if (debug_enabled())
            fprintf(stderr, "[gzgpu] %s: %zu members in one pipeline, %s%s\n", what, mem.size(),
                    bad == mem.size() ? "all ok" : "first failure: ",
                    bad == mem.size() ? "" : (mem[bad].msg.empty() ? "not at its span's end" : mem[bad].msg.c_str()));
        if (bad != mem.size()) {
            c->st = gzgpu_stats{};
            return false;
        }
        sink_rc = drain_to_sink(c, c->d_out.as<uint8_t>(), out_total, sink);
        *total = out_total;
        c->st.compressed_bytes = in_len;
        c->st.decompressed_bytes = out_total;
        if (stats) *stats = c->st;
        return true;
    };
    int sink_rc = 0;
    std::function<int(const uint8_t *, size_t)> sink_fn = [&](const uint8_t *d, size_t n) { return sink(d, n); };
    const std::vector<uint64_t> moff = sized_member_walk(in, in_len);
    if (moff.size() > 2) {
        std::vector<std::pair<size_t, size_t>> sp;
        for (size_t i = 0; i + 1 < moff.size(); i++) sp.push_back({moff[i], moff[i + 1] - moff[i]});
        if (together(sp, false, "sized members", sink_rc)) return sink_rc;
    }
    const std::vector<std::pair<size_t, size_t>> spans = member_spans(in, in_len);
    if (moff.size() <= 2 && spans.size() > 1) {
        bool plausible = true;
        for (auto &sp : spans) plausible = plausible && span_isize(in + sp.first, sp.second) != ~0ull;
        if (plausible && together(spans, true, "candidate members", sink_rc)) return sink_rc;
    }
    auto next_candidate = [&](size_t off) -> size_t {
        auto it = std::upper_bound(spans.begin(), spans.end(), off,
                                   [](size_t v, const std::pair<size_t, size_t> &sp) { return v < sp.first; });
        return it == spans.end() ? in_len : it->first;
    };
    size_t off = 0;
    int members = 0;
    while (off < in_len) {
        if (members > 0 && (in_len - off < 2 || in[off] != 0x1f || in[off + 1] != 0x8b)) break;  // trailing bytes
        // first the span to the next candidate start, then (if the member
        // does not end there) the rest of the input
        const size_t bound = next_candidate(off);
        Member M;
        bool done = false;
        for (int attempt = bound < in_len ? 0 : 1; attempt < 2 && !done; attempt++) {
            const size_t avail = (attempt == 0 ? bound : in_len) - off;
            std::string msg;
            int st = member_from_host(in + off, avail, off, M, msg);
            if (st) {
                if (attempt == 0) continue;
                if (members > 0) {
                    off = in_len;  // trailing bytes that are not a member
                    break;
                }
                return finish(c, st, msg);
            }
            M.split = true;
            // the member may be followed by others: size the output from
            // the trailer at the end of the span, growing on overflow
            uint64_t cap = std::max<uint64_t>((uint64_t)M.exp_isize + 64 * KIB, 1 * MIB);
            // the input's only member: its output streams to the sink while
            // the decode finishes (resolve ranges, D2H on a copy stream)
            const bool stream = attempt == 1 && members == 0 && spans.size() <= 1 && moff.size() <= 2;
            for (int tries = 0;; tries++) {
                c->d_out.ensure(cap);
                M.out = c->d_out.as<uint8_t>();
                M.out_cap = c->d_out.cap;
                M.status = GZGPU_OK;
                std::vector<Member> v{M};
                c->streamed = false;
                c->stream_rc = 0;
                c->stream_sink = stream ? &sink_fn : nullptr;
                try {
                    run_members(c, c->d_in.as<uint8_t>(), in_len, v);
                } catch (...) {
                    c->stream_sink = nullptr;
                    throw;
                }
                c->stream_sink = nullptr;
                M = v[0];
                if (M.status == GZGPU_OUTPUT_TOO_SMALL && tries < 8 && !c->streamed) {
                    cap = c->d_out.cap * 4;
                    continue;
                }
                break;
            }
            if (c->streamed && c->stream_rc) return c->stream_rc;
            if (attempt == 0) {
                // accepted only when it ends exactly at this span's trailer
                done = M.status == GZGPU_OK && (M.end_bit + 7) / 8 == M.de / 8;
                continue;
            }
            // the member's real end: trailer right after the final block
            if (M.status == GZGPU_OK || M.status == GZGPU_SIZE_MISMATCH || M.status == GZGPU_CRC_MISMATCH) {
                uint64_t tr = (M.end_bit + 7) / 8;
                if (tr + 8 <= in_len && tr + 8 != in_len) {
                    // more bytes follow: re-check against this member's own trailer
                    uint32_t crc = le32(in + tr), isz = le32(in + tr + 4);
                    M.exp_crc = crc;
                    M.exp_isize = isz;
                    if ((uint32_t)M.out_len != isz) {
                        M.status = GZGPU_SIZE_MISMATCH;
                        M.msg = "ISIZE mismatch";
                    } else if (M.crc != crc) {
                        M.status = GZGPU_CRC_MISMATCH;
                        M.msg = "CRC32 mismatch";
                    } else {
                        M.status = GZGPU_OK;
                    }
                } else if (tr + 8 > in_len) {
                    M.status = GZGPU_DECODE_FAILED;
                    M.msg = "truncated gzip trailer";
                }
            }
            if (M.status != GZGPU_OK) return finish(c, M.status, M.msg);
            done = true;
        }
        if (!done) break;
        if (!c->streamed) {
            int sr = drain_to_sink(c, M.out, M.out_len, sink);
            if (sr) return sr;
        }
        c->streamed = false;
        *total += M.out_len;
        off = (M.end_bit + 7) / 8 + 8;
        members++;
    }
    c->st.compressed_bytes = in_len;
    c->st.decompressed_bytes = *total;
    if (stats) *stats = c->st;
    return GZGPU_OK;
}

}  // namespace

// ---- scratch of the calls without allocation ------------------------------
// One array per device, held for the life of the process.  Calls that
// overlap on one device share it and race: callers serialise them.
namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = 256ull << 20;  // device buffers of one call
constexpr size_t NOALLOC_STAGE_BYTES = 8ull << 20;      // pinned host staging per context
constexpr uint32_t NOALLOC_MAX_PIECES = 256;
}  // namespace

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
// The array's address on `device` (the current device), looked up once per
// device.  The name `scratch` in host code is a host-side stand-in, not a
// device address: host code uses only this pointer.
uint8_t *device_scratch(int device) {
    static std::mutex mu;
    static std::unordered_map<int, uint8_t *> ptrs;
    std::lock_guard<std::mutex> g(mu);
    auto it = ptrs.find(device);
    if (it != ptrs.end()) return it->second;
    uint8_t *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw StatusError{GZGPU_GPU_UNAVAILABLE, "no address for the scratch array"};
    ptrs[device] = ptr;
    return ptr;
}
}  // namespace

// =========================================================================
// C ABI
// =========================================================================
extern "C" {

int gzgpu_ctx_create(int device, gzgpu_ctx **out) {
    if (!out) return GZGPU_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) {
        (void)hipGetLastError();
        return GZGPU_GPU_UNAVAILABLE;
    }
    gzgpu_ctx *c = new gzgpu_ctx();
    c->device = device;
    int st = guarded(c, [&]() {
        HIPCHK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw StatusError{GZGPU_GPU_UNAVAILABLE, std::string("device is ") + prop.gcnArchName + ", need gfx950"};
        c->num_cus = prop.multiProcessorCount;
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
        for (auto &e : c->res_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        init_constants(c);
        {
            // workgroups of one wavefront: blocks per CU = wavefronts per CU
            int occ[4] = {0, 0, 0, 0};
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[0], k_inflate<uint8_t, 9>, 64, 0));
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[1], k_inflate<uint8_t, 10>, 64, 0));
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[2], k_inflate<uint16_t, 9>, 64, 0));
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[3], k_inflate<uint16_t, 10>, 64, 0));
            // LDS goes to workgroups in 1280-byte granules (gfx950's 160 KiB
            // is 128 of them): at 11952 bytes (the 10-bit root) 12 fit, not
            // the 13 the query reports (measured: the 13th wavefront of
            // each CU started only when another had finished)
            const uint32_t lds[4] = {(uint32_t)sizeof(InflateShared<uint8_t, 9>),
                                     (uint32_t)sizeof(InflateShared<uint8_t, 10>),
                                     (uint32_t)sizeof(InflateShared<uint16_t, 9>),
                                     (uint32_t)sizeof(InflateShared<uint16_t, 10>)};
            for (int i = 0; i < 4; i++) {
                const uint32_t by_lds = (160u * 1024u) / (uint32_t)round_up(lds[i], 1280);
                c->inflate_occ[i / 2][i % 2] = std::max(1u, std::min({16u, (uint32_t)std::max(occ[i], 1), by_lds}));
            }
            if (debug_enabled())
                fprintf(stderr, "[gzgpu] inflate wavefronts per CU: u8 %u/%u, u16 %u/%u (9/10-bit root)\n",
                        c->inflate_occ[0][0], c->inflate_occ[0][1], c->inflate_occ[1][0], c->inflate_occ[1][1]);
        }
        c->scratch = device_scratch(device);
        HIPCHK(hipHostMalloc((void **)&c->noalloc_stage, NOALLOC_STAGE_BYTES, hipHostMallocDefault));
        return (int)GZGPU_OK;
    });
    if (st != GZGPU_OK) {
        fprintf(stderr, "gzgpu_ctx_create: %s\n", c->err.c_str());
        gzgpu_ctx_destroy(c);
        return st;
    }
    *out = c;
    return GZGPU_OK;
}

void gzgpu_ctx_destroy(gzgpu_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (DevBuf *b : ctx_bufs(c)) b->release();
    if (c->noalloc_stage) (void)hipHostFree(c->noalloc_stage);
    for (auto &b : c->extra_slabs) b.release();
    c->pin.release();
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c->res_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    for (int i = 0; i < 2; i++) {
        if (c->drain_ev[i]) (void)hipEventDestroy(c->drain_ev[i]);
        if (c->drain[i]) (void)hipHostFree(c->drain[i]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int gzgpu_set_tuning(gzgpu_ctx *c, const gzgpu_tuning *t) {
    if (!c || !t) return GZGPU_INVALID_ARGUMENT;
    if (t->lut_bits != 0 && t->lut_bits != 9 && t->lut_bits != 10) {
        set_err(c, "lut_bits must be 0 (automatic), 9 or 10");
        return GZGPU_INVALID_ARGUMENT;
    }
    c->tune = *t;
    return GZGPU_OK;
}

int gzgpu_get_tuning(gzgpu_ctx *c, gzgpu_tuning *t) {
    if (!c || !t) return GZGPU_INVALID_ARGUMENT;
    *t = c->tune;  // as set: 0 fields stay 0 (automatic)
    return GZGPU_OK;
}

const char *gzgpu_last_error(gzgpu_ctx *c) { return c ? c->err.c_str() : "no context"; }

const char *gzgpu_status_string(int s) {
    switch (s) {
    case GZGPU_OK: return "ok";
    case GZGPU_INVALID_HEADER: return "invalid gzip header";
    case GZGPU_INVALID_FORMAT: return "input below parallel SM minimum (routing bug)";
    case GZGPU_DECODE_FAILED: return "chunk decode failed";
    case GZGPU_SIZE_MISMATCH: return "output size mismatch";
    case GZGPU_CRC_MISMATCH: return "CRC32 mismatch";
    case GZGPU_IO: return "I/O error";
    case GZGPU_GPU_UNAVAILABLE: return "GPU unavailable";
    case GZGPU_OUT_OF_DEVICE_MEMORY: return "out of device memory";
    case GZGPU_OUTPUT_TOO_SMALL: return "output buffer too small";
    case GZGPU_INVALID_ARGUMENT: return "invalid argument";
    case GZGPU_PENDING: return "request pending";
    default: return "unknown status";
    }
}

int gzgpu_decompress(gzgpu_ctx *c, const uint8_t *in, size_t in_len, uint8_t *out, size_t out_cap, size_t *out_len,
                     gzgpu_stats *stats) {
    if (!c || (!in && in_len) || !out_len) return GZGPU_INVALID_ARGUMENT;
    return guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        size_t written = 0;
        bool too_small = false;
        auto sink = [&](const uint8_t *d, size_t n) -> int {
            if (!too_small && written + n <= out_cap) {
                if (n) std::memcpy(out + written, d, n);
            } else {
                too_small = true;
            }
            written += n;
            return 0;
        };
        uint64_t total = 0;
        int st = decompress_host_stream(c, in, in_len, sink, &total, stats);
        *out_len = written;
        if (st) return st;
        if (too_small) {
            set_err(c, "output buffer too small");
            return (int)GZGPU_OUTPUT_TOO_SMALL;
        }
        return (int)GZGPU_OK;
    });
}

int gzgpu_decompress_to_writer(gzgpu_ctx *c, const uint8_t *in, size_t in_len, gzgpu_write_fn write, void *user,
                               uint64_t *total, gzgpu_stats *stats) {
    if (!c || (!in && in_len) || !write || !total) return GZGPU_INVALID_ARGUMENT;
    return guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        auto sink = [&](const uint8_t *d, size_t n) -> int {
            if (!n) return 0;
            if (write(user, d, n) != 0) {
                set_err(c, "writer callback failed");
                return (int)GZGPU_IO;
            }
            return 0;
        };
        return decompress_host_stream(c, in, in_len, sink, total, stats);
    });
}

}  // extern "C"

namespace {

// One member, device-resident in and out.  max_isize: members whose trailer
// ISIZE exceeds it are refused before any device work (calls without
// allocation).
int decompress_device_member(gzgpu_ctx *c, const void *d_in, size_t in_len, void *d_out, size_t out_cap,
                             size_t *out_len, uint32_t *crc_out, gzgpu_stats *stats, uint64_t max_isize) {
    {
        HIPCHK(hipSetDevice(c->device));
        c->st = gzgpu_stats{};
        *out_len = 0;
        // envelope: header from the first bytes, trailer from the last 8,
        // into pinned staging (one small transfer each, one wait); a header
        // longer than the first 4 KiB (FEXTRA / FNAME / FCOMMENT) is read
        // again in full below
        size_t head = std::min<size_t>(in_len, 4 * KIB);
        uint8_t *h = (uint8_t *)c->pin.alloc(head + 8);
        memset(h + head, 0, 8);
        HIPCHK(hipMemcpyAsync(h, d_in, head, hipMemcpyDeviceToHost, c->stream));
        if (in_len >= 8)
            HIPCHK(hipMemcpyAsync(h + head, (const uint8_t *)d_in + in_len - 8, 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (in_len < 2 || h[0] != 0x1f || h[1] != 0x8b) return (int)GZGPU_OK;
        size_t hl;
        int st = parse_header(h, head, &hl);
        const uint32_t exp_crc = le32(h + head), exp_isize = le32(h + head + 4);
        // a long FEXTRA / FNAME / FCOMMENT: read more of the header
        for (size_t more = head; st == GZGPU_INVALID_HEADER && more < in_len;) {
            more = std::min<size_t>(in_len, 4 * more);
            std::vector<uint8_t> big(more);
            HIPCHK(hipMemcpyAsync(big.data(), d_in, more, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            st = parse_header(big.data(), more, &hl);
        }
        if (st) return finish(c, st, "invalid gzip header");
        if (in_len < hl + 8) return finish(c, GZGPU_INVALID_FORMAT, "input shorter than header + trailer");
        if (exp_isize > max_isize)
            return finish(c, GZGPU_INVALID_ARGUMENT,
                          "member of " + std::to_string(exp_isize) + " bytes (trailer ISIZE) exceeds the limit of " +
                              std::to_string(max_isize));
        Member M;
        M.ds = hl * 8;
        M.de = (in_len - 8) * 8;
        M.limit = in_len * 8;
        M.exp_crc = exp_crc;
        M.exp_isize = exp_isize;
        M.out = (uint8_t *)d_out;
        M.out_cap = out_cap;
        M.split = true;
        std::vector<Member> v{M};
        run_members(c, (const uint8_t *)d_in, in_len, v);
        M = v[0];
        *out_len = M.out_len;
        if (crc_out) *crc_out = M.crc;
        c->st.compressed_bytes = in_len;
        c->st.decompressed_bytes = M.out_len;
        if (stats) *stats = c->st;
        return finish(c, M.status, M.msg);
    }
}

}  // namespace

extern "C" {

int gzgpu_decompress_device(gzgpu_ctx *c, const void *d_in, size_t in_len, void *d_out, size_t out_cap,
                            size_t *out_len, uint32_t *crc_out, gzgpu_stats *stats) {
    if (!c || !d_in || !d_out || !out_len) return GZGPU_INVALID_ARGUMENT;
    return guarded(c, [&]() {
        return decompress_device_member(c, d_in, in_len, d_out, out_cap, out_len, crc_out, stats, ~0ull);
    });
}

// The same decode with no allocation: every device buffer comes from this
// device's `scratch` array, host staging from the pinned block made with the
// context; the planner caps the pieces so that a member within the limits
// fits.  The context's own buffers are set aside for the call and put back
// after it, untouched.
int gzgpu_decompress_device_noalloc(gzgpu_ctx *c, const void *d_in, size_t in_len, void *d_out, size_t out_cap,
                                    size_t *out_len, uint32_t *crc_out, gzgpu_stats *stats) {
    if (!c || !d_in || !d_out || !out_len) return GZGPU_INVALID_ARGUMENT;
    *out_len = 0;
    if (!c->scratch || !c->noalloc_stage) {
        set_err(c, "no scratch memory on this device");
        return GZGPU_GPU_UNAVAILABLE;
    }
    if (in_len > GZGPU_NOALLOC_MAX_IN_BYTES) {
        set_err(c, "input of " + std::to_string(in_len) + " bytes exceeds the no-allocation limit of " +
                       std::to_string((uint64_t)GZGPU_NOALLOC_MAX_IN_BYTES));
        return GZGPU_INVALID_ARGUMENT;
    }
    std::vector<DevBuf> saved;
    const std::vector<DevBuf *> bufs = ctx_bufs(c);
    saved.reserve(bufs.size());
    for (DevBuf *b : bufs) {
        saved.push_back(*b);
        *b = DevBuf{};
    }
    ScratchArena arena{c->scratch, NOALLOC_SCRATCH_BYTES, 0};
    const gzgpu_tuning tune = c->tune;
    c->tune = gzgpu_tuning{};
    c->max_pieces = NOALLOC_MAX_PIECES;
    c->pin.fixed = c->noalloc_stage;
    c->pin.fixed_cap = NOALLOC_STAGE_BYTES;
    c->pin.fixed_used = 0;
    t_arena = &arena;
    const int st = guarded(
        c,
        [&]() {
            return decompress_device_member(c, d_in, in_len, d_out, out_cap, out_len, crc_out, stats,
                                            GZGPU_NOALLOC_MAX_OUT_BYTES);
        },
        false);
    // the stream is idle here (every path syncs before returning or
    // throwing out of the pipeline, and guarded waits at entry); wait anyway
    // before the scratch is handed to the next call
    (void)hipStreamSynchronize(c->stream);
    t_arena = nullptr;
    c->pin.fixed = nullptr;
    c->max_pieces = 0;
    c->tune = tune;
    for (size_t i = 0; i < bufs.size(); i++) *bufs[i] = saved[i];
    for (auto &b : c->extra_slabs) b.release();
    c->extra_slabs.clear();
    return st;
}

int gzgpu_decompress_batch(gzgpu_ctx *c, const void *d_in, const uint64_t *in_off, const uint64_t *in_len, void *d_out,
                           const uint64_t *out_off, const uint64_t *out_cap, size_t n, int32_t *status,
                           uint64_t *out_len, uint32_t *crc, gzgpu_stats *stats) {
    if (!c || (!n) || !d_in || !d_out || !in_off || !in_len || !out_off || !out_cap) return GZGPU_INVALID_ARGUMENT;
    return guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        c->st = gzgpu_stats{};
        gzgpu_tuning tn;
        gzgpu_get_tuning(c, &tn);
        // gather every member's first bytes and trailer in one launch
        const size_t per = ENVELOPE_HEAD + 8;
        upload(c, c->d_env_off, in_off, n * 8);
        upload(c, c->d_env_len, in_len, n * 8);
        c->d_env.ensure(n * per);
        hipLaunchKernelGGL(k_gather_envelope, dim3((uint32_t)std::min<size_t>((n * per + 255) / 256, 65535)), dim3(256), 0,
                           c->stream, (const uint8_t *)d_in, c->d_env_off.as<uint64_t>(), c->d_env_len.as<uint64_t>(),
                           (uint32_t)n, c->d_env.as<uint8_t>());
        HIPCHK(hipGetLastError());
        std::vector<uint8_t> env(n * per);
        HIPCHK(hipMemcpyAsync(env.data(), c->d_env.p, env.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        uint64_t in_total = 0;
        for (size_t i = 0; i < n; i++) in_total = std::max<uint64_t>(in_total, in_off[i] + in_len[i]);
        std::vector<Member> mem(n);
        for (size_t i = 0; i < n; i++) {
            Member &M = mem[i];
            const uint8_t *e = env.data() + i * per;
            size_t hl = 0;
            int st = parse_header(e, (size_t)std::min<uint64_t>(in_len[i], ENVELOPE_HEAD), &hl);
            if (st == GZGPU_INVALID_HEADER && in_len[i] > ENVELOPE_HEAD) {
                // long FEXTRA/FNAME: fetch the full header
                std::vector<uint8_t> big((size_t)std::min<uint64_t>(in_len[i], 1 * MIB));
                HIPCHK(hipMemcpy(big.data(), (const uint8_t *)d_in + in_off[i], big.size(), hipMemcpyDeviceToHost));
                st = parse_header(big.data(), big.size(), &hl);
            }
            if (st || in_len[i] < hl + 8) {
                M.status = st ? st : GZGPU_INVALID_FORMAT;
                M.msg = "invalid gzip header";
                continue;
            }
            M.ds = (in_off[i] + hl) * 8;
            M.de = (in_off[i] + in_len[i] - 8) * 8;
            M.limit = (in_off[i] + in_len[i]) * 8;
            M.exp_crc = le32(e + ENVELOPE_HEAD);
            M.exp_isize = le32(e + ENVELOPE_HEAD + 4);
            M.out = (uint8_t *)d_out + out_off[i];
            M.out_cap = out_cap[i];
            // 0 (default): the planner cuts members by expected output
            M.split = tn.batch_split_bytes == 0 || in_len[i] > tn.batch_split_bytes;
        }
        run_members(c, (const uint8_t *)d_in, in_total, mem);
        int first_bad = GZGPU_OK;
        for (size_t i = 0; i < n; i++) {
            if (status) status[i] = mem[i].status;
            if (out_len) out_len[i] = mem[i].out_len;
            if (crc) crc[i] = mem[i].crc;
            if (mem[i].status != GZGPU_OK && first_bad == GZGPU_OK) {
                first_bad = mem[i].status;
                set_err(c, "member " + std::to_string(i) + ": " + mem[i].msg);
            }
        }
        c->st.decompressed_bytes = 0;
        for (auto &M : mem) c->st.decompressed_bytes += M.out_len;
        for (size_t i = 0; i < n; i++) c->st.compressed_bytes += in_len[i];
        if (stats) *stats = c->st;
        return first_bad;
    });
}

int gzgpu_autotune(gzgpu_ctx *c, const void *d_in, const uint64_t *in_off, const uint64_t *in_len, void *d_out,
                   const uint64_t *out_off, const uint64_t *out_cap, size_t n, uint32_t budget_ms,
                   gzgpu_tuning *best, double *best_ms) {
    if (!c || !n || !d_in || !d_out || !in_off || !in_len || !out_off || !out_cap) return GZGPU_INVALID_ARGUMENT;
    // candidates: the planner's own grid (piece per resident slot, by decode
    // cost) at two occupancies, then explicit grids; batch split rules
    std::vector<gzgpu_tuning> cand;
    for (uint64_t split : {(uint64_t)0, (uint64_t)(1 * MIB), (uint64_t)(4 * MIB)}) {
        cand.push_back({0, 0, 0, split, 0, 0});
        if (split) continue;
        cand.push_back({0, 0, 0, split, 0, 10});  // a 10-bit litlen LUT root (13 wavefronts per CU)
        cand.push_back({0, 0, 3, split, 0, 0});
        for (uint64_t chunk : {256 * KIB, 512 * KIB, 1024 * KIB})
            for (uint32_t ppc : {4u, 8u, 16u}) cand.push_back({chunk, ppc, 0, split, 0, 0});
    }
    const gzgpu_tuning keep = c->tune;
    std::vector<int32_t> status(n);
    std::vector<uint64_t> olen(n);
    std::vector<uint32_t> crc(n);
    // Each candidate: one run to warm its buffers, then the best of two.  The
    // planner's default (the first candidate) is kept unless another beats it
    // by more than AUTOTUNE_MARGIN: on the 10,000-member batch the default
    // and three wavefronts per SIMD measured within noise of each other, and
    // a single timed run picked either from one process to the next (the
    // batch then ran at 94 or 102 GB/s).
    constexpr double AUTOTUNE_MARGIN = 0.02;
    double best_t = 1e30, spent = 0, t_default = 1e30;
    gzgpu_tuning pick = keep;
    for (auto &t : cand) {
        if (spent > (double)budget_ms && best_t < 1e30) break;
        c->tune = t;
        double tmin = 1e30;
        for (int rep = 0; rep < 3; rep++) {  // first run warms the buffers
            gzgpu_stats st{};
            int r = gzgpu_decompress_batch(c, d_in, in_off, in_len, d_out, out_off, out_cap, n, status.data(),
                                           olen.data(), crc.data(), &st);
            if (r != GZGPU_OK) {
                c->tune = keep;
                return r;
            }
            spent += st.ms_total;
            if (rep) tmin = std::min(tmin, st.ms_total);
        }
        if (&t == &cand[0]) t_default = tmin;
        if (tmin < best_t) {
            best_t = tmin;
            pick = t;
        }
    }
    if (best_t > (1.0 - AUTOTUNE_MARGIN) * t_default) {
        pick = cand[0];
        best_t = t_default;
    }
    c->tune = pick;
    if (best) *best = pick;
    if (best_ms) *best_ms = best_t;
    return GZGPU_OK;
}

int gzgpu_find_blocks(gzgpu_ctx *c, const void *d_in, size_t in_len, uint64_t start_bit, uint64_t end_bit,
                      gzgpu_boundary *out, size_t cap, size_t *count) {
    if (!c || !d_in || !count) return GZGPU_INVALID_ARGUMENT;
    return guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        uint64_t eb = std::min<uint64_t>(end_bit, (uint64_t)in_len * 8);
        const AlignedIn ai = align_input(d_in);
        const uint64_t sb = ai.shift * 8;
        std::vector<gzgpu_boundary> h =
            eb > start_bit ? find_hits(c, ai.base, in_len + ai.shift, start_bit + sb, eb + sb)
                           : std::vector<gzgpu_boundary>();
        for (auto &b : h) {
            b.bit_offset -= sb;
            b.seek_bit -= sb;
            if (b.kind == 2) b.info &= 0xFFFFFu;  // the planner's output estimate stays internal
        }
        *count = h.size();
        if (out) std::memcpy(out, h.data(), std::min(cap, h.size()) * sizeof(gzgpu_boundary));
        return (int)GZGPU_OK;
    });
}

int gzgpu_crc32(gzgpu_ctx *c, const void *d_data, size_t len, uint32_t *crc) {
    if (!c || (!d_data && len) || !crc) return GZGPU_INVALID_ARGUMENT;
    return guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        ensure_crc_tables(c);
        std::vector<uint64_t> so, st;
        std::vector<uint32_t> sl;
        // segment boundaries on 16-byte addresses
        const uint64_t lead = (uintptr_t)d_data & 15;
        for (uint64_t o = 0; o < len;) {
            const uint64_t l = std::min<uint64_t>(o ? CRC_SEG : CRC_SEG - lead, len - o);
            so.push_back(o);
            sl.push_back((uint32_t)l);
            st.push_back(len - o - l);
            o += l;
        }
        uint32_t nseg = (uint32_t)so.size();
        c->d_seg_crc.ensure(std::max<size_t>(nseg, 1) * 4);
        if (nseg) {
            upload(c, c->d_seg_off, so.data(), so.size() * 8);
            upload(c, c->d_seg_len, sl.data(), sl.size() * 4);
            upload(c, c->d_seg_tail, st.data(), st.size() * 8);
            hipLaunchKernelGGL(k_crc_segments, dim3(std::min<uint32_t>(nseg, (uint32_t)c->num_cus * 8)), dim3(256), 0,
                               c->stream, (const uint8_t *)d_data, c->d_seg_off.as<uint64_t>(),
                               c->d_seg_len.as<uint32_t>(), c->d_seg_tail.as<uint64_t>(), nseg,
                               c->d_crc_tab.as<uint32_t>(), c->d_seg_crc.as<uint32_t>());
            HIPCHK(hipGetLastError());
        }
        uint32_t ms[2] = {0, nseg};
        uint64_t ml = len;
        upload(c, c->d_mseg, ms, 8);
        upload(c, c->d_mlen, &ml, 8);
        c->d_mcrc.ensure(4);
        hipLaunchKernelGGL(k_crc_members, dim3(1), dim3(256), 0, c->stream, c->d_seg_crc.as<uint32_t>(),
                           c->d_mseg.as<uint32_t>(), c->d_mlen.as<uint64_t>(), 1u, c->d_mcrc.as<uint32_t>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(crc, c->d_mcrc.p, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return (int)GZGPU_OK;
    });
}

int gzgpu_inflate_piece(gzgpu_ctx *c, const void *d_in, size_t in_len, uint64_t start_bit, uint64_t stop_bit,
                        const uint8_t *window, void *d_out, size_t out_cap_elems, gzgpu_piece_result *res) {
    if (!c || !d_in || !d_out || !res) return GZGPU_INVALID_ARGUMENT;
    return guarded(c, [&]() {
        HIPCHK(hipSetDevice(c->device));
        const AlignedIn ai = align_input(d_in);
        const uint64_t sb = ai.shift * 8;
        PieceDesc pd{};
        pd.start_bit = start_bit + sb;
        pd.stop_bit = stop_bit >= NO_STOP - sb ? NO_STOP : stop_bit + sb;
        pd.limit_bit = ((uint64_t)in_len + ai.shift) * 8;
        pd.member = 0;
        DevBuf slab, win;
        if (window) {
            // clean decode against a known window, straight into d_out
            win.ensure(WIN);
            HIPCHK(hipMemcpyAsync(win.p, window, WIN, hipMemcpyHostToDevice, c->stream));
            pd.window = win.as<uint8_t>();
            pd.dst = (uint8_t *)d_out;
            pd.cap = out_cap_elems;
            pd.flags = PF_WINDOW;
        } else {
            uint64_t cap = marker_slab_bytes(out_cap_elems);
            slab.ensure(cap);
            pd.dst = slab.as<uint8_t>();
            pd.cap = cap;
            pd.flags = PF_MARKER;
        }
        std::vector<PieceDesc> d1{pd};
        std::vector<uint64_t> cost{1};
        float ms = 0;
        const uint64_t n_in = (uint64_t)in_len + ai.shift;
        const PieceOut po = window ? finish_inflate(c, enqueue_inflate<uint8_t>(c, ai.base, n_in, d1, cost), d1, ms)[0]
                                   : finish_inflate(c, enqueue_inflate<uint16_t>(c, ai.base, n_in, d1, cost), d1, ms)[0];
        res->end_bit = po.end_bit >= sb ? po.end_bit - sb : 0;
        res->bfinal_hit = po.status == PS_FINAL;
        res->status = po.status == PS_FINAL || po.status == PS_STOPPED ? GZGPU_OK
                      : po.status == PS_OVERFLOW                      ? GZGPU_OUTPUT_TOO_SMALL
                                                                      : GZGPU_DECODE_FAILED;
        uint64_t n = po.n16 + po.n8;
        res->out_len = n;
        if (!window) {
            // u16 view of every entry: marker part as is, clean part widened
            if (n > out_cap_elems) n = out_cap_elems;
            uint64_t a = std::min<uint64_t>(po.n16, n);
            HIPCHK(hipMemcpyAsync(d_out, pd.dst, a * 2, hipMemcpyDeviceToDevice, c->stream));
            if (n > a) {
                std::vector<uint8_t> tmp(n - a);
                HIPCHK(hipMemcpyAsync(tmp.data(), pd.dst + po.u8_off, n - a, hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipStreamSynchronize(c->stream));
                std::vector<uint16_t> w(tmp.begin(), tmp.end());
                HIPCHK(hipMemcpyAsync((uint16_t *)d_out + a, w.data(), w.size() * 2, hipMemcpyHostToDevice, c->stream));
            }
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        slab.release();
        win.release();
        return (int)GZGPU_OK;
    });
}

}  // extern "C"

#include "serve.h"
#include "graph.h"
#include "index.h"
