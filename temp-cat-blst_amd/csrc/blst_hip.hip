// Host side of the gfx950 G1 MSM backend: per-device context, launch
// sequence of the device stages, the latency-bound window fold, and the C ABI
// declared in include/blst_hip.h.
//
// The drop-in entry points keep the reference's checks and their order
// (blst/c_api.cpp:1382-1418, 1546-1611): they are the only part of the
// reference's control flow that callers can observe.  Everything past the
// checks is this backend's own pipeline.
#include "pipeline.h"
#include "host_ec2.h"
#include "../../include/blst_hip.h"

#include <dlfcn.h>
#include <condition_variable>
#include <exception>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

using namespace bls;

namespace bls {
namespace host {

thread_local std::string g_err;
thread_local double g_phase[7];
std::atomic<size_t> g_alloc_cap{0};
std::atomic<int> g_stage_fail_range{-1};

// the message of the exception being handled (called inside a catch block):
// no exception crosses the C ABI, every one becomes blst_hip_last_error()
std::string current_error() {
    try {
        throw;
    } catch (const HipFail& f) {
        return f.msg;
    } catch (const std::exception& e) {
        return e.what();
    } catch (...) {
        return "unknown error";
    }
}

// Parallel host memcpy, for staging pageable caller buffers into page-locked
// memory: copy() splits the bytes over the pool's threads and the caller.
class CopyPool {
public:
    explicit CopyPool(unsigned n) {
        for (unsigned i = 0; i + 1 < n; i++) workers_.emplace_back([this] { loop(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    void copy(void* dst, const void* src, size_t bytes) {
        const size_t parts = workers_.size() + 1;
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<uint8_t*>(dst);
            src_ = static_cast<const uint8_t*>(src);
            bytes_ = bytes;
            step_ = std::max<size_t>(((bytes + parts - 1) / parts + 4095) & ~(size_t)4095, 4096);
            next_.store(0);
            pending_ = workers_.size();
            gen_++;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    void drain() {
        for (size_t i; (i = next_.fetch_add(1)) * step_ < bytes_;) {
            const size_t off = i * step_;
            std::memcpy(dst_ + off, src_ + off, std::min(step_, bytes_ - off));
        }
    }
    void loop() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            lk.unlock();
            drain();
            lk.lock();
            if (--pending_ == 0) done_cv_.notify_all();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    uint8_t* dst_ = nullptr;
    const uint8_t* src_ = nullptr;
    size_t bytes_ = 0, step_ = 1;
    std::atomic<size_t> next_{0};
    size_t pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Host-to-device uploads of caller buffers.  Page-locked or device memory
// goes straight to the copy engines; a large pageable buffer is copied by
// CopyPool threads into a ring of page-locked chunks, each chunk's DMA
// running while the next chunk is filled (at 2^20 points this took the
// host-buffer MSM from 10.5 to 8.0 ms; the runtime's own pageable path moved
// about 21 GB/s).
struct Stager {
    static constexpr size_t kChunk = (size_t)16 << 20;
    static constexpr int kRing = 3;
    static constexpr size_t kMin = (size_t)8 << 20;   // smaller copies go direct
    uint8_t* ring[kRing] = {};
    hipEvent_t ev[kRing] = {};
    uint64_t next_slot = 0;   // round robin over the ring across calls
    std::unique_ptr<CopyPool> pool;
    ~Stager() {
        for (int k = 0; k < kRing; k++) {
            if (ring[k]) (void)hipHostFree(ring[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
        }
    }
    void init() {
        if (pool) return;
        for (int k = 0; k < kRing; k++) {
            HIPCK(hipHostMalloc((void**)&ring[k], kChunk, hipHostMallocDefault));
            HIPCK(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        }
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        pool = std::make_unique<CopyPool>(std::max(1u, std::min(8u, hw / 2)));
    }
    // page-locked host or device memory (the copy engines reach it directly)
    static bool registered(const void* p) {
        hipPointerAttribute_t a;
        const bool r = hipPointerGetAttributes(&a, p) == hipSuccess && a.type != hipMemoryTypeUnregistered;
        (void)hipGetLastError();
        return r;
    }
    void upload(void* dst, const void* src, size_t bytes, hipStream_t stream) {
        const bool direct = bytes < kMin || registered(src);
        if (direct) {
            HIPCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
            return;
        }
        init();
        const uint8_t* s = static_cast<const uint8_t*>(src);
        uint8_t* d = static_cast<uint8_t*>(dst);
        // the slots continue from the previous call's, so the first chunk of a
        // call does not wait for the DMA of the previous call's first chunk
        for (size_t off = 0; off < bytes; off += kChunk) {
            const int slot = (int)(next_slot++ % kRing);
            HIPCK(hipEventSynchronize(ev[slot]));   // the slot's previous DMA is done
            const size_t len = std::min(kChunk, bytes - off);
            pool->copy(ring[slot], s + off, len);
            HIPCK(hipMemcpyAsync(d + off, ring[slot], len, hipMemcpyHostToDevice, stream));
            HIPCK(hipEventRecord(ev[slot], stream));
        }
    }
    // device-to-host after the work queued on stream; returns when dst holds
    // the bytes: a pageable dst gets the chunks from the ring, each copied out
    // by the pool while the next chunks' DMAs run
    void download(void* dst, const void* src, size_t bytes, hipStream_t stream) {
        if (bytes < kMin || registered(dst)) {
            HIPCK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
            HIPCK(hipStreamSynchronize(stream));
            return;
        }
        init();
        const uint8_t* s = static_cast<const uint8_t*>(src);
        uint8_t* d = static_cast<uint8_t*>(dst);
        const size_t nch = (bytes + kChunk - 1) / kChunk;
        auto len = [&](size_t k) { return std::min(kChunk, bytes - k * kChunk); };
        auto issue = [&](size_t k) {
            const int slot = (int)(k % kRing);
            HIPCK(hipMemcpyAsync(ring[slot], s + k * kChunk, len(k), hipMemcpyDeviceToHost, stream));
            HIPCK(hipEventRecord(ev[slot], stream));
        };
        for (size_t k = 0; k < std::min<size_t>(nch, kRing); k++) issue(k);
        for (size_t k = 0; k < nch; k++) {
            const int slot = (int)(k % kRing);
            HIPCK(hipEventSynchronize(ev[slot]));
            pool->copy(d + k * kChunk, ring[slot], len(k));
            if (k + kRing < nch) issue(k + kRing);
        }
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;        // == ws.stream
    std::mutex mu;
    Workspace ws;
    std::vector<std::unique_ptr<Workspace>> pipe;   // window-group pipelines
    hipEvent_t ready = nullptr;                     // inputs written on stream
    std::vector<hipEvent_t> range_ready;            // per point range of a split host MSM
    DevBuf pts, scl;
    DevBuf jac_in, aff_out;              // batch normalisation staging
    DevBuf wb_pts, wb_rows, wb_table, wb_list, wb_scl, wb_part, wb_W;   // fixed-base tables
    DevBuf comb_q, comb_x;               // comb-table build: shifted bases
    void* wb_host = nullptr;             // page-locked gather staging
    size_t wb_host_cap = 0;
    Stager stager;                       // uploads of pageable caller buffers
    Stager down;                         // downloads that run beside uploads
    hipStream_t d2h = nullptr;           // their copies (created on first use)
};

std::mutex g_ctx_mu;
std::map<int, std::unique_ptr<Ctx>> g_ctx;

Ctx& context() {
    int dev = 0;
    HIPCK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_ctx.find(dev);
    if (it != g_ctx.end()) return *it->second;
    auto c = std::make_unique<Ctx>();
    c->device = dev;
    // The context's stream and the two window-group pipelines' streams are
    // created together, first: HIP maps streams onto a few hardware queues
    // (GPU_MAX_HW_QUEUES, four by default) in creation order, and two
    // pipelines whose streams share a queue run one after the other (a 2^20
    // MSM then took 5.0 ms instead of 4.06).
    c->ws.init();
    c->stream = c->ws.stream;
    for (int g = 0; g < 2; g++) {
        c->pipe.emplace_back(new Workspace());
        c->pipe.back()->init();
    }
    HIPCK(hipEventCreateWithFlags(&c->ready, hipEventDisableTiming));
    Ctx& ref = *c;
    g_ctx.emplace(dev, std::move(c));
    return ref;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---------------------------------------------------------------- planning

// The reference's window heuristic (pippenger_window_size,
// blst/scalar/curve/multi_scalar.h:329-348); it fixes the scratch sizes the
// drop-in entry points report.
size_t ref_window_size(size_t npoints) {
    size_t wbits;
    for (wbits = 0; npoints >>= 1; wbits++) {}
    if (wbits > 12) wbits -= 3;
    else if (wbits > 8) wbits -= 2;
    else if (wbits > 4) wbits -= 1;
    else return wbits ? 2 : 1;
    return wbits < 24 ? wbits : 24;
}

// Device window: minimises n*D + k*2^(c-1)*D, the count of bucket additions
// plus the bucket-reduction additions weighted by their relative cost.
uint32_t device_window(size_t n, size_t nbits) {
    double best = 1e300;
    uint32_t bc = 1;
    for (uint32_t c = 1; c <= 20; c++) {
        const double D = 1.0 + (double)(nbits / c);
        const double cost = (double)n * D + 6.0 * (double)(1u << (c - 1)) * D;
        if (cost < best) { best = cost; bc = c; }
    }
    return bc;
}

// ------------------------------------------------------------------ tuning
//
// Two tables (a lone MSM, and a job inside a batch) indexed by floor(log2 n),
// filled from the autotuner's file when present: BLST_HIP_TUNING, else
// tuning.txt next to this library.  Lines: "single|batch log2n c T L" for
// full-width scalars (200 bits or more), and "single64|single128 log2n c T L"
// for lone MSMs of 33..64-bit and 65..128-bit scalars.  Entries the file
// does not set fall back to the cost model.

struct TuneTable {
    Tune t[32];
    bool set[32];
};

std::mutex g_tune_mu;
TuneTable g_tune[2];
TuneTable g_tune_short[2];   // lone MSMs of 33..64-bit and 65..128-bit scalars
bool g_tune_loaded = false;

// the short-scalar table of a width (-1: none)
int short_class(size_t nbits) { return nbits > 32 && nbits <= 64 ? 0 : (nbits > 64 && nbits <= 128 ? 1 : -1); }

void load_tuning_locked() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    for (auto& tb : g_tune)
        for (bool& b : tb.set) b = false;
    for (auto& tb : g_tune_short)
        for (bool& b : tb.set) b = false;
    std::string path;
    if (const char* env = std::getenv("BLST_HIP_TUNING")) {
        path = env;
    } else {
        Dl_info info;
        if (dladdr((void*)&load_tuning_locked, &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning.txt";
        }
    }
    FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
    if (!f) return;
    char mode[16];
    unsigned lg, c, T, L;
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        if (std::sscanf(line, "%15s %u %u %u %u", mode, &lg, &c, &T, &L) != 5) continue;
        const int m = std::strcmp(mode, "batch") == 0      ? 1
                      : std::strcmp(mode, "single") == 0   ? 0
                      : std::strcmp(mode, "single64") == 0  ? 2
                      : std::strcmp(mode, "single128") == 0 ? 3
                                                            : -1;
        if (m < 0 || lg > 31 || c < 1 || c > 24 || T < 1 || T > 4096 || L < 1 || L > 4096) continue;
        TuneTable& tb = m < 2 ? g_tune[m] : g_tune_short[m - 2];
        tb.t[lg] = Tune{c, T, L};
        tb.set[lg] = true;
    }
    std::fclose(f);
}

int floor_log2(size_t n) {
    int l = 0;
    while (n >>= 1) l++;
    return l;
}

Tune tuned(size_t n, size_t nbits, bool batch) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_tuning_locked();
    const int lg = std::min(floor_log2(std::max<size_t>(n, 1)), 31);
    // the main tables are measured for full-width scalars
    if (nbits >= 200 && g_tune[batch ? 1 : 0].set[lg]) return g_tune[batch ? 1 : 0].t[lg];
    const int sc = short_class(nbits);
    if (!batch && sc >= 0 && g_tune_short[sc].set[lg]) return g_tune_short[sc].t[lg];
    return Tune{device_window(n, nbits), 32, 8};
}

void set_tuning(bool batch, int lg, Tune t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_tuning_locked();
    g_tune[batch ? 1 : 0].t[lg] = t;
    g_tune[batch ? 1 : 0].set[lg] = true;
}

// job description with the tuned (or explicit) window and tile shape
JobSpec make_spec(const void* points, const void* scalars, size_t n, size_t nbits, size_t window, bool batch) {
    const Tune t = tuned(n, nbits, batch);
    JobSpec js;
    js.points = points;
    js.scalars = scalars;
    js.n = (uint32_t)n;
    js.nbits = (uint32_t)nbits;
    js.c = window ? (uint32_t)window : t.c;
    js.T = t.T;
    js.L = t.L;
    js.bit_base = 0;
    js.tile = false;
    return js;
}

Layout make_layout(const std::vector<JobSpec>& specs) {
    Layout L;
    L.g2 = !specs.empty() && specs[0].g2;
    uint64_t pts = 0, ent = 0, keys = 0, segs = 0, biggest = 0, tiles = 0;
    for (const JobSpec& s : specs) {
        if (s.g2 != L.g2) throw std::runtime_error("a batch mixes G1 and G2 jobs");
        job_t J;
        J.scalars = static_cast<const uint8_t*>(s.scalars);
        J.points = static_cast<const uint4*>(s.points);
        J.n = s.n;
        J.nbits = s.nbits;
        J.nbytes = (s.nbits + 7) / 8;
        J.c = s.c;
        J.nwin = s.tile ? 1u : (s.nwin ? s.nwin : 1u + s.nbits / s.c);
        J.B = 1u << (s.c - 1);
        J.bit_base = s.tile ? s.bit_base : s.win0 * s.c;
        J.vec = (J.nbytes == 32 && ((uintptr_t)s.scalars % 16) == 0) ? 1u : 0u;
        J.kbase = (uint32_t)keys;
        J.ebase = (uint32_t)ent;
        J.pbase = (uint32_t)pts;
        J.gbase = (uint32_t)L.groups.size();
        for (uint32_t w = 0; w < J.nwin; w++) {
            group_t G = {};
            G.kb = J.kbase + w * J.B;
            G.B = J.B;
            uint32_t seg = 1;
            while (seg * 2 <= std::max<uint32_t>(s.L, 1)) seg *= 2;   // a power of two divides B
            G.L = std::min<uint32_t>(seg, J.B);
            G.P = J.B / G.L;
            G.sb = (uint32_t)segs;
            segs += G.P;
            L.groups.push_back(G);
            L.sbase.push_back(G.sb);
            // sort tiles: chunks of max(2B, 4096) points per window
            sgroup_t S = {};
            S.job = (uint32_t)L.jobs.size();
            S.w = w;
            S.ch = std::max<uint32_t>(2 * J.B, 4096);
            S.nch = std::max<uint32_t>(div_up(J.n, S.ch), 1);
            S.cb = 0;
            tiles += S.nch;
            L.sgroups.push_back(S);
            L.gkb.push_back(G.kb);
        }
        L.max_B = std::max(L.max_B, J.B);
        if ((uint64_t)J.nwin * J.n > biggest) {
            biggest = (uint64_t)J.nwin * J.n;
            L.T = std::max<uint32_t>(s.T, 1);
        }
        pts += J.n;
        ent += (uint64_t)J.nwin * J.n;
        keys += (uint64_t)J.nwin * J.B;
        L.max_n = std::max(L.max_n, J.n);
        L.jobs.push_back(J);
        L.pbase.push_back(J.pbase);
        L.ebase.push_back(J.ebase);
        L.kbase.push_back(J.kbase);
    }
    if (ent >= 0x7fffffffULL || keys >= 0x7fffffffULL || pts >= 0x7fffffffULL || segs >= 0x7fffffffULL)
        throw HipFail{"batch too large for one launch"};
    if (tiles >= 0x7fffffffULL) throw HipFail{"batch too large for one launch"};
    // the binned sort when every window has 2^6 .. 2^15 buckets (window key
    // ranges then start at multiples of the bin size) and at least half of
    // the entries fall into windows of 2^12 buckets or more, whose one-pass
    // runs would hold one or two entries per bucket and tile
    {
        bool fits = !L.groups.empty();
        uint64_t wide = 0, all = 0;
        for (const sgroup_t& S : L.sgroups) {
            const job_t& J = L.jobs[S.job];
            fits = fits && J.B >= (1u << kBinBits) && J.B <= (kBinMaxBins << kBinBits);
            all += J.n;
            wide += J.B >= (1u << 12) ? J.n : 0;
        }
        L.binned = fits && 2 * wide >= all;
    }
    if (L.binned) {
        size_t cw = 0, nt = 0;
        for (sgroup_t& S : L.sgroups) {
            const job_t& J = L.jobs[S.job];
            S.ch = kBinTile;
            S.nch = std::max<uint32_t>(div_up(J.n, kBinTile), 1);
            S.cb = (uint32_t)cw;
            cw += (size_t)S.nch * (J.B >> kBinBits);
            nt += S.nch;
        }
        if (cw >= 0x7fffffffULL || nt >= 0x7fffffffULL) throw HipFail{"batch too large for one launch"};
        L.count_words = cw;
        L.ntiles = (uint32_t)nt;
        L.tiles.reserve(nt);
        for (size_t g = 0; g < L.sgroups.size(); g++)
            for (uint32_t k = 0; k < L.sgroups[g].nch; k++) L.tiles.push_back(make_uint2((uint32_t)g, k));
        L.tile_begin = {0u, (uint32_t)nt};
        L.tile_lds = {0};
    }
    size_t cw = 0;
    for (sgroup_t& S : L.sgroups) {
        if (L.binned) break;
        S.cb = (uint32_t)cw;
        cw += (size_t)S.nch * L.jobs[S.job].B;
    }
    if (cw >= 0xffffffffULL) throw HipFail{"batch too large for one launch"};
    if (!L.binned) {
        L.count_words = cw;
        L.ntiles = (uint32_t)tiles;
    }
    // tile table by LDS size class: a launch reserves the LDS of its widest
    // window for every tile, so narrow windows of a mixed batch get their own
    // launches and share CUs
    if (!L.binned) {
        const uint32_t limits[4] = {1u << 9, 1u << 12, 1u << 14, 0xffffffffu};
        L.tiles.reserve(tiles);
        L.tile_begin.push_back(0);
        for (int cls = 0; cls < 4; cls++) {
            uint32_t widest = 0;
            for (size_t g = 0; g < L.sgroups.size(); g++) {
                const uint32_t B = L.groups[g].B;
                const uint32_t lo = cls ? limits[cls - 1] : 0;
                if (B <= lo || B > limits[cls]) continue;
                widest = std::max(widest, B);
                for (uint32_t k = 0; k < L.sgroups[g].nch; k++) L.tiles.push_back(make_uint2((uint32_t)g, k));
            }
            if ((uint32_t)L.tiles.size() == L.tile_begin.back()) continue;
            L.tile_begin.push_back((uint32_t)L.tiles.size());
            L.tile_lds.push_back((size_t)widest * 4);
        }
    }
    // reduction tree passes: each block merges up to kTreeNodes nodes of a
    // group into one, until one node per group is left
    {
        const size_t ng = L.groups.size();
        std::vector<uint32_t> n(ng), in(ng);
        for (size_t g = 0; g < ng; g++) {
            n[g] = L.groups[g].P;
            in[g] = L.groups[g].sb;
        }
        size_t in_off = 0, in_nodes = segs;
        while (true) {
            uint32_t blocks = 0, outn = 0;
            bool more = false;
            for (size_t g = 0; g < ng; g++) {
                const uint32_t b = div_up(n[g], kTreeNodes);
                L.tree_blk.push_back(blocks);
                L.tree_tn.push_back(make_uint4(in[g], n[g], outn, 0));
                blocks += b;
                in[g] = outn;
                outn += b;
                n[g] = b;
                more = more || b > 1;
            }
            L.tree_blocks.push_back(blocks);
            L.tree_in.push_back(in_off);
            L.tree_out.push_back(in_off + in_nodes);
            if (!more) break;
            in_off += in_nodes;
            in_nodes = outn;
        }
        L.tree_nodes = L.tree_out.back();
    }
    L.total_points = (uint32_t)pts;
    L.total_entries = (uint32_t)ent;
    L.nkeys = (uint32_t)keys;
    L.total_segments = (uint32_t)segs;
    // a dispatch counts its work-items in 32 bits: every launch of the batch
    // must stay below 2^32 of them (many jobs of many narrow windows reach it
    // first in the sort, one 512-thread workgroup per (job, window) tile)
    {
        const uint64_t nchunks = std::max<uint64_t>((ent + L.T - 1) / L.T, 1);
        const uint64_t merge_blocks = 2 * ((nchunks + 255) / 256) + 256 + 512;
        uint64_t items = (uint64_t)L.ntiles * kSortThreads;
        items = std::max<uint64_t>(items, merge_blocks * 256);
        if (L.binned) items = std::max<uint64_t>(items, (keys >> kBinBits) * 256);
        if (L.g2) items = std::max<uint64_t>(items, 4 * segs);
        for (uint32_t b : L.tree_blocks) items = std::max<uint64_t>(items, (uint64_t)b * 256);
        if (items > 0xffffffffULL) throw HipFail{"batch too large for one launch"};
    }
    return L;
}

// bytes (a multiple of 4) between device memory and a page-locked host
// buffer on cx.stream: a copy kernel through the host buffer's device mapping
// when the workspace asks for kernel transfers, else an async copy
void transfer_to(Workspace& cx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    if (cx.kernel_transfers) {
        void* mapped = nullptr;
        const void* host = kind == hipMemcpyHostToDevice ? src : dst;
        if (hipHostGetDevicePointer(&mapped, const_cast<void*>(host), 0) == hipSuccess && mapped) {
            const uint32_t* s = static_cast<const uint32_t*>(kind == hipMemcpyHostToDevice ? mapped : src);
            uint32_t* d = static_cast<uint32_t*>(kind == hipMemcpyHostToDevice ? dst : mapped);
            launch_k_copy_words(cx.stream, s, d, bytes / 4);
            HIPCK(hipGetLastError());
            return;
        }
        (void)hipGetLastError();
    }
    HIPCK(hipMemcpyAsync(dst, src, bytes, kind, cx.stream));
}

size_t layout_table_bytes(const Layout& L) {
    const size_t nj = L.jobs.size(), ng = L.groups.size();
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t t = al(nj * sizeof(job_t));
    t = al(t + nj * 4);
    t = al(t + nj * 4);
    t = al(t + nj * 4);
    t = al(t + ng * sizeof(group_t));
    t = al(t + ng * 4);
    t = al(t + ng * sizeof(sgroup_t));
    t = al(t + L.tiles.size() * sizeof(uint2));
    t = al(t + ng * 4);
    t = al(t + L.tree_blk.size() * 4);
    return al(t + L.tree_tn.size() * sizeof(uint4));
}

DevLayout upload_layout(Workspace& cx, const Layout& L, size_t base) {
    const size_t nj = L.jobs.size(), ng = L.groups.size();
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_jobs = 0, o_p = al(o_jobs + nj * sizeof(job_t)), o_e = al(o_p + nj * 4),
                 o_k = al(o_e + nj * 4), o_g = al(o_k + nj * 4), o_s = al(o_g + ng * sizeof(group_t)),
                 o_sg = al(o_s + ng * 4), o_tb = al(o_sg + ng * sizeof(sgroup_t)),
                 o_gk = al(o_tb + L.tiles.size() * sizeof(uint2)),
                 o_trb = al(o_gk + ng * 4), o_trt = al(o_trb + L.tree_blk.size() * 4),
                 total = al(o_trt + L.tree_tn.size() * sizeof(uint4));
    cx.tables.ensure(base + total);
    uint8_t* h = cx.tables_stage(base + total) + base;
    std::memcpy(h + o_jobs, L.jobs.data(), nj * sizeof(job_t));
    std::memcpy(h + o_p, L.pbase.data(), nj * 4);
    std::memcpy(h + o_e, L.ebase.data(), nj * 4);
    std::memcpy(h + o_k, L.kbase.data(), nj * 4);
    std::memcpy(h + o_g, L.groups.data(), ng * sizeof(group_t));
    std::memcpy(h + o_s, L.sbase.data(), ng * 4);
    std::memcpy(h + o_sg, L.sgroups.data(), ng * sizeof(sgroup_t));
    std::memcpy(h + o_tb, L.tiles.data(), L.tiles.size() * sizeof(uint2));
    std::memcpy(h + o_gk, L.gkb.data(), ng * 4);
    std::memcpy(h + o_trb, L.tree_blk.data(), L.tree_blk.size() * 4);
    std::memcpy(h + o_trt, L.tree_tn.data(), L.tree_tn.size() * sizeof(uint4));
    uint8_t* d = cx.tables.as<uint8_t>() + base;
    transfer_to(cx, d, h, total, hipMemcpyHostToDevice);
    DevLayout D;
    D.jobs = reinterpret_cast<const job_t*>(d + o_jobs);
    D.pbase = reinterpret_cast<const uint32_t*>(d + o_p);
    D.ebase = reinterpret_cast<const uint32_t*>(d + o_e);
    D.kbase = reinterpret_cast<const uint32_t*>(d + o_k);
    D.groups = reinterpret_cast<const group_t*>(d + o_g);
    D.sbase = reinterpret_cast<const uint32_t*>(d + o_s);
    D.sgroups = reinterpret_cast<const sgroup_t*>(d + o_sg);
    D.tiles = reinterpret_cast<const uint2*>(d + o_tb);
    D.gkb = reinterpret_cast<const uint32_t*>(d + o_gk);
    D.tree_blk = reinterpret_cast<const uint32_t*>(d + o_trb);
    D.tree_tn = reinterpret_cast<const uint4*>(d + o_trt);
    return D;
}

// ---------------------------------------------------------------- device

void scan_exclusive(Workspace& cx, const uint32_t* in, uint32_t* out, uint32_t N, uint32_t* tmp) {
    const uint32_t tiles = div_up(N, kScanTile);
    if (tiles == 1) {
        launch_k_scan_tiles(dim3(1), dim3(kScanThreads), cx.stream, in, out, nullptr, N);
        return;
    }
    uint32_t* sums = tmp;
    uint32_t* sums_scanned = tmp + tiles;
    launch_k_scan_tiles(dim3(tiles), dim3(kScanThreads), cx.stream, in, out, sums, N);
    scan_exclusive(cx, sums, sums_scanned, tiles, tmp + 2 * (size_t)tiles);
    launch_k_scan_add(dim3(div_up(N, 256)), dim3(256), cx.stream, out, sums_scanned, N);
}

size_t scan_tmp_words(uint32_t N) {
    size_t w = 0;
    while (true) {
        const uint32_t tiles = div_up(N, kScanTile);
        if (tiles == 1) break;
        w += 2 * (size_t)tiles;
        N = tiles;
    }
    return w + 16;
}

void fill_words(Workspace& cx, void* dst, size_t words, uint32_t value) {
    const uint32_t blocks = (uint32_t)std::min<size_t>(div_up(words, 256), 8192);
    launch_k_fill(dim3(std::max<uint32_t>(blocks, 1)), dim3(256), cx.stream,
                       static_cast<uint32_t*>(dst), words, value);
}

// The tiled sort keeps a window's bucket counters in LDS; wider windows take
// the global-histogram path.
bool tiled_sort(const Layout& L) {
    static const size_t limit = sort_lds_limit();
    return (size_t)L.max_B * 4 <= limit;
}

// Every device buffer of a batch's workspace with the bytes the layout needs
// (f(buffer, bytes)); the layout-table buffer is sized by upload_layout.
template <class Fn>
void for_each_buffer(Workspace& cx, const Layout& L, Fn&& f) {
    const uint32_t nkeys = L.nkeys, nent = L.total_entries;
    const uint32_t nslots = 2 * std::max<uint32_t>(div_up(nent, L.T), 1);
    if (L.binned) {
        f(cx.counts, (L.count_words + 1) * 4 + 16);
        f(cx.rank, (L.count_words + 1) * 4 + 16);   // scanned bin counts
        f(cx.ent, (size_t)nent * 8 + 16);           // entries in bin order
    } else if (tiled_sort(L)) {
        f(cx.counts, L.count_words * 4 + 16);
    } else {
        f(cx.ent, (size_t)nent * 4 + 16);
        f(cx.rank, (size_t)nent * 4 + 16);
    }
    f(cx.sorted, (size_t)nent * 8 + 16);
    f(cx.hist, ((size_t)nkeys + 1) * 4);
    f(cx.off, ((size_t)nkeys + 1) * 4);
    f(cx.scan_tmp, std::max(scan_tmp_words(nkeys + 1), scan_tmp_words((uint32_t)L.count_words + 1)) * 4);
    const size_t pb = L.point_bytes();
    f(cx.buckets, (size_t)nkeys * pb);
    f(cx.slot_key, (size_t)nslots * 4);
    f(cx.slots, (size_t)nslots * pb);
    f(cx.heavy, (3 * (size_t)(nslots / 2) + kListHeadWords) * 4);   // list lengths, part counters, three lists
    f(cx.H, L.tree_nodes * 2 * pb);
    f(cx.W, (size_t)L.groups.size() * pb);
    f(cx.res, (size_t)L.jobs.size() * pb);
}

void ensure_workspace(Workspace& cx, const Layout& L) {
    for_each_buffer(cx, L, [](DevBuf& b, size_t bytes) { b.ensure(bytes); });
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t carved_bytes(const Layout& L) {
    Workspace probe;   // only its buffer members are named; nothing is created
    size_t off = 0;
    for_each_buffer(probe, L, [&](DevBuf&, size_t bytes) { off = align256(off) + bytes; });
    return align256(off) + layout_table_bytes(L);
}

void carve_workspace(Workspace& cx, const Layout& L, uint8_t* base, size_t cap) {
    const size_t need = carved_bytes(L);
    if (need > cap)
        throw HipFail{"MSM needs " + std::to_string(need) + " bytes of scratch, the fixed region holds " +
                      std::to_string(cap)};
    size_t off = 0;
    for_each_buffer(cx, L, [&](DevBuf& b, size_t bytes) {
        off = align256(off);
        b.attach(base + off, bytes);
        off += bytes;
    });
    off = align256(off);
    cx.tables.attach(base + off, cap - off);
}

// Device bytes a batch of this layout needs beyond its inputs.
size_t workspace_bytes(const Layout& L) {
    const size_t nslots = 2 * (size_t)std::max<uint32_t>(div_up(L.total_entries, L.T), 1);
    const size_t sort_bytes = L.binned ? (L.count_words + 1) * 8 + (size_t)L.total_entries * 8
                              : tiled_sort(L) ? L.count_words * 4 : (size_t)L.total_entries * 8;
    const size_t pb = L.point_bytes();
    return (size_t)L.total_entries * 8 + sort_bytes + ((size_t)L.nkeys + 1) * 8 + (size_t)L.nkeys * pb +
           nslots * (10 + pb) + L.tree_nodes * 2 * pb + (L.groups.size() + L.jobs.size()) * pb + (1 << 20);
}

// The point stages of a batch (accumulate, partial merges, reduction, fold)
// over the field of its group: F = fp for G1, fp2 for G2.
template <class F>
void enqueue_points(Workspace& cx, const Layout& L, const DevLayout& D, bool host_fold, void* host_out,
                    bool timing, const RangeRole& range) {
    using P = xyzz_t<F>;
    const uint32_t nkeys = L.nkeys;
    const uint32_t nent = L.total_entries;
    const uint32_t T = L.T;
    const uint32_t nchunks = std::max<uint32_t>(div_up(nent, T), 1);
    const uint32_t ngroups = (uint32_t)L.groups.size();
    const uint32_t njobs = (uint32_t)L.jobs.size();
    hipStream_t st = cx.stream;
    uint32_t* hist = cx.hist.as<uint32_t>();
    uint32_t* off = cx.off.as<uint32_t>();
    // buckets are not cleared: every bucket with entries is written below,
    // and the reduction skips the empty ones by their count.  The
    // accumulation lists the buckets of more than two pieces for the merge
    // (list lengths and the heavy merge's part counters cleared first).
    uint32_t* heavy = cx.heavy.as<uint32_t>();
    merge_lists_t lists;
    lists.counts = heavy;
    lists.parts = heavy + 64;
    lists.wave = heavy + kListHeadWords;
    lists.group = lists.wave + nchunks;
    lists.few = lists.group + nchunks;
    fill_words(cx, heavy, kListHeadWords, 0);
    if (timing) HIPCK(hipEventRecord(cx.ev[2], st));
    launch_k_accumulate(dim3(div_up(nchunks, 256)), dim3(256), st, cx.sorted.as<const uint2>(), off + nkeys, D.jobs,
                        D.kbase, njobs, cx.buckets.as<P>(), cx.slot_key.as<uint32_t>(), cx.slots.as<P>(), T,
                        nchunks, range.prev, off, hist, lists);
    HIPCK(hipGetLastError());
    if (timing) HIPCK(hipEventRecord(cx.ev[3], st));

    // buckets spanning chunks: every kind of merge in one launch
    launch_k_partial_merge(st, cx.slot_key.as<const uint32_t>(), cx.slots.as<P>(), off, hist, T, nchunks,
                           cx.buckets.as<P>(), lists);
    HIPCK(hipGetLastError());
    if (timing) HIPCK(hipEventRecord(cx.ev[4], st));
    if (range.cnt_sum) launch_k_add_words(st, hist, range.cnt_sum, nkeys, range.prev == nullptr);
    if (!range.reduce) return;
    const uint32_t* cnt = range.cnt_sum ? range.cnt_sum : hist;

    P* Hn = cx.H.as<P>();
    launch_k_reduce_segments(dim3(div_up(L.total_segments, 256)), dim3(256), st, D.groups, D.sbase, ngroups,
                             L.total_segments, cx.buckets.as<const P>(), cnt, Hn);
    const size_t npass = L.tree_blocks.size();
    for (size_t p = 0; p < npass; p++)
        launch_k_reduce_tree(dim3(L.tree_blocks[p]), st, D.tree_blk + p * ngroups, D.tree_tn + p * ngroups, ngroups,
                             Hn + 2 * L.tree_in[p], Hn + 2 * L.tree_out[p], cx.W.as<P>(),
                             p + 1 == npass ? 1u : 0u);
    HIPCK(hipGetLastError());
    if (host_fold) {
        transfer_to(cx, host_out, cx.W.p, (size_t)ngroups * sizeof(P), hipMemcpyDeviceToHost);
    } else {
        launch_k_fold(st, D.jobs, njobs, cx.W.as<const P>(), cx.res.as<P>());
        HIPCK(hipGetLastError());
        transfer_to(cx, host_out, cx.res.p, (size_t)njobs * sizeof(P), hipMemcpyDeviceToHost);
    }
    if (timing) HIPCK(hipEventRecord(cx.ev[5], st));
}

// Enqueues the device pipeline of a batch on cx.stream; nothing in it waits
// for the host, so the sequence can be captured in a HIP graph.  With host_fold the window sums W[g] are copied to host_out (one xyzz per
// group); otherwise k_fold combines the windows of every job and one xyzz per
// job is copied.  host_out should be pinned for the copy to be asynchronous.
void enqueue_batch(Workspace& cx, const Layout& L, const DevLayout& D, bool host_fold, void* host_out,
                   bool timing, const RangeRole& range) {
    const uint32_t nkeys = L.nkeys;
    const uint32_t nent = L.total_entries;
    const uint32_t ngroups = (uint32_t)L.groups.size();
    const uint32_t njobs = (uint32_t)L.jobs.size();
    hipStream_t st = cx.stream;
    uint32_t* hist = cx.hist.as<uint32_t>();
    uint32_t* off = cx.off.as<uint32_t>();

    if (timing) HIPCK(hipEventRecord(cx.ev[0], st));
    if (L.binned) {
        uint32_t* counts = cx.counts.as<uint32_t>();
        uint32_t* pos = cx.rank.as<uint32_t>();
        uint2* tmp = cx.ent.as<uint2>();
        const uint32_t nc = (uint32_t)L.count_words;
        launch_k_bin_count(dim3(L.ntiles), st, D.jobs, D.sgroups, D.tiles, counts, nc);
        HIPCK(hipGetLastError());
        if (timing) HIPCK(hipEventRecord(cx.ev[1], st));
        scan_exclusive(cx, counts, pos, nc + 1, cx.scan_tmp.as<uint32_t>());
        launch_k_bin_place(dim3(L.ntiles), st, D.jobs, D.sgroups, D.tiles, pos, tmp);
        launch_k_bin_fine(dim3(nkeys >> kBinBits), st, D.sgroups, D.gkb, ngroups, pos, tmp, cx.sorted.as<uint2>(),
                          hist, off, nkeys, nc);
    } else if (tiled_sort(L)) {
        uint32_t* counts = cx.counts.as<uint32_t>();
        for (size_t c = 0; c < L.tile_lds.size(); c++)
            launch_k_count(dim3(L.tile_begin[c + 1] - L.tile_begin[c]), dim3(kSortThreads), L.tile_lds[c], st, D.jobs,
                           D.sgroups, D.tiles + L.tile_begin[c], counts);
        launch_k_prefix(dim3(div_up(nkeys + 1, 256)), dim3(256), st, D.groups, D.sgroups, D.gkb, ngroups, nkeys,
                        counts, hist);
        HIPCK(hipGetLastError());
        if (timing) HIPCK(hipEventRecord(cx.ev[1], st));
        scan_exclusive(cx, hist, off, nkeys + 1, cx.scan_tmp.as<uint32_t>());
        for (size_t c = 0; c < L.tile_lds.size(); c++)
            launch_k_place(dim3(L.tile_begin[c + 1] - L.tile_begin[c]), dim3(kSortThreads), L.tile_lds[c], st, D.jobs,
                           D.sgroups, D.tiles + L.tile_begin[c], counts, off, cx.sorted.as<uint2>());
    } else {
        fill_words(cx, hist, (size_t)nkeys + 1, 0);
        launch_k_digits(dim3(div_up(L.total_points, 256)), dim3(256), st, D.jobs, D.pbase, njobs, L.total_points,
                        cx.ent.as<uint32_t>(), cx.rank.as<uint32_t>(), hist);
        HIPCK(hipGetLastError());
        if (timing) HIPCK(hipEventRecord(cx.ev[1], st));
        scan_exclusive(cx, hist, off, nkeys + 1, cx.scan_tmp.as<uint32_t>());
        launch_k_scatter(dim3(div_up(nent, 256)), dim3(256), st, D.jobs, D.ebase, njobs, nent,
                         cx.ent.as<const uint32_t>(), cx.rank.as<const uint32_t>(), off, cx.sorted.as<uint2>());
    }
    HIPCK(hipGetLastError());
    if (L.g2) enqueue_points<fp2>(cx, L, D, host_fold, host_out, timing, range);
    else enqueue_points<fp>(cx, L, D, host_fold, host_out, timing, range);
}

void record_phases(Workspace& cx) {
    float ms;
    for (int k = 0; k < 5; k++) {
        HIPCK(hipEventElapsedTime(&ms, cx.ev[k], cx.ev[k + 1]));
        g_phase[k] = ms;
    }
}

// Synchronous batch: layout upload, pipeline, results in out.
void run_batch(Workspace& cx, const Layout& L, bool host_fold, std::vector<hostec::xyzz6>& out) {
    ensure_workspace(cx, L);
    const DevLayout D = upload_layout(cx, L);
    out.resize(host_fold ? L.groups.size() : L.jobs.size());
    enqueue_batch(cx, L, D, host_fold, out.data());
    HIPCK(hipStreamSynchronize(cx.stream));
    record_phases(cx);
}

// window sums of one job -> blst_p1
void finish_single(const JobSpec& js, const std::vector<hostec::xyzz6>& wins, blst_p1* ret) {
    auto t0 = std::chrono::steady_clock::now();
    hostec::xyzz6 r;
    if (!js.tile) hostec::fold_windows(r, wins.data(), (int)wins.size(), (int)js.c);
    else r = wins[0];
    hostec::jac6 j;
    hostec::to_jac(j, r);
    static_assert(sizeof(hostec::jac6) == sizeof(blst_p1), "blst_p1 layout");
    std::memcpy(ret, &j, sizeof(j));
    g_phase[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Window groups of one MSM run as separate pipelines on their own streams,
// so the sort of one group overlaps the accumulation of another and the
// latency-bound merges and reduction of a group overlap the accumulation of
// the next.  0 = automatic.
std::atomic<uint32_t> g_pipeline{0};
// percent of the windows in the last group (0: equal groups)
std::atomic<uint32_t> g_last_share{0};

// First window of group g of G over nwin windows; group 0 holds the top
// windows and group G - 1 the lowest.  The last group's results are the ones
// no other work hides (its merges, reduction and host fold run after every
// accumulation), so it may be given fewer windows than the others.
uint32_t group_win0(uint32_t g, uint32_t G, uint32_t nwin) {
    if (g + 1 >= G) return 0;
    const uint32_t pct = g_last_share.load();
    uint32_t last = pct ? (nwin * pct + 50) / 100 : nwin / G;
    last = std::max<uint32_t>(1, std::min<uint32_t>(last, nwin - (G - 1)));
    const uint32_t rest = nwin - last;
    return last + (G - 2 - g) * rest / (G - 1);
}

uint32_t group_win_end(uint32_t g, uint32_t G, uint32_t nwin) { return g == 0 ? nwin : group_win0(g - 1, G, nwin); }

uint32_t pipeline_groups(const JobSpec& js) {
    if (js.tile) return 1;
    const uint32_t nwin = 1 + js.nbits / js.c;
    uint32_t G = g_pipeline.load();
    if (G == 0) G = js.n >= (1u << 17) ? 2 : 1;
    return std::max<uint32_t>(1, std::min<uint32_t>(G, nwin / 2));
}

// The window sums of G window-group pipelines (group g on cx.pipe[g], the
// top windows in group 0, outs[g] its nsums[g] sums) folded on the host into
// ret.  Horner over the groups from the top: acc holds the groups above g at
// the base of group g - 1; its shift down to group g's base needs no data of
// group g and runs before waiting for it.
void fold_groups(Ctx& cx, uint32_t G, uint32_t nwin, uint32_t c, const std::vector<hostec::xyzz6*>& outs,
                 const std::vector<uint32_t>& nsums, blst_p1* ret) {
    hostec::xyzz6 acc;
    hostec::set_inf(acc);
    std::chrono::steady_clock::time_point t0;
    for (uint32_t g = 0; g < G; g++) {
        const uint32_t base = group_win0(g, G, nwin);
        if (g > 0) hostec::scale_pow2(acc, acc, (int)(c * (group_win_end(g, G, nwin) - base)));
        HIPCK(hipStreamSynchronize(cx.pipe[g]->stream));
        if (g + 1 == G) t0 = std::chrono::steady_clock::now();
        hostec::xyzz6 part;
        hostec::fold_windows(part, outs[g], (int)nsums[g], (int)c);
        hostec::addp(acc, acc, part);
    }
    record_phases(*cx.pipe[0]);
    hostec::jac6 j;
    hostec::to_jac(j, acc);
    std::memcpy(ret, &j, sizeof(j));
    // the fold time left after the last group's results arrive
    g_phase[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// one MSM (or tile) over device-resident inputs written on cx.stream
void device_msm(Ctx& cx, const JobSpec& js, blst_p1* ret) {
    const uint32_t G = pipeline_groups(js);
    if (G == 1) {
        std::vector<JobSpec> v(1, js);
        Layout L = make_layout(v);
        std::vector<hostec::xyzz6> wins;
        run_batch(cx.ws, L, true, wins);
        finish_single(js, wins, ret);
        return;
    }
    const uint32_t nwin = 1 + js.nbits / js.c;
    while (cx.pipe.size() < G) {
        cx.pipe.emplace_back(new Workspace());
        cx.pipe.back()->init();
    }
    // layouts, buffers and tables first (allocation may synchronise), then
    // the pipelines, each behind the inputs
    std::vector<Layout> Ls(G);
    std::vector<DevLayout> Ds(G);
    std::vector<hostec::xyzz6*> outs(G);
    // group 0 holds the top windows: it is queued first and finishes first,
    // and the host folds it while the lower groups still run
    for (uint32_t g = 0; g < G; g++) {
        JobSpec s = js;
        s.win0 = group_win0(g, G, nwin);
        s.nwin = group_win_end(g, G, nwin) - s.win0;
        Ls[g] = make_layout(std::vector<JobSpec>(1, s));
        Workspace& w = *cx.pipe[g];
        ensure_workspace(w, Ls[g]);
        Ds[g] = upload_layout(w, Ls[g]);
        outs[g] = static_cast<hostec::xyzz6*>(w.host_buf(Ls[g].groups.size() * sizeof(hostec::xyzz6)));
    }
    HIPCK(hipEventRecord(cx.ready, cx.stream));
    for (uint32_t g = 0; g < G; g++) {
        Workspace& w = *cx.pipe[g];
        HIPCK(hipStreamWaitEvent(w.stream, cx.ready, 0));
        enqueue_batch(w, Ls[g], Ds[g], true, outs[g], g == 0);
    }
    std::vector<uint32_t> nsums(G);
    for (uint32_t g = 0; g < G; g++) nsums[g] = (uint32_t)Ls[g].groups.size();
    fold_groups(cx, G, nwin, js.c, outs, nsums, ret);
}

// The staging of a split MSM's point ranges on a thread of its own:
// stage(r, lo, m) (host copies into the page-locked ring and their DMAs, on
// cx.stream) for r = 0 .. K-1 in turn, each followed by cx.range_ready[r]
// recorded on cx.stream.  The caller's thread queues the kernels of range r
// once wait(r) returns, so the copies of the later ranges run while the
// earlier ranges' kernels are being queued (on one thread the two took turns:
// a range's copies began only after the previous range's launches).
// wait(r) and join() rethrow what stopped the staging; the destructor only
// joins (a caller unwinding from its own error waits for the copies queued).
class RangeStager {
public:
    template <class Stage>
    RangeStager(Ctx& cx, const std::vector<size_t>& lo, Stage& stage) : K_((uint32_t)lo.size() - 1) {
        thread_ = std::thread([this, &cx, &lo, &stage] {
            uint32_t r = 0;
            try {
                HIPCK(hipSetDevice(cx.device));
                for (; r < K_; r++) {
                    if ((int)r == g_stage_fail_range.load(std::memory_order_relaxed))
                        throw HipFail{"staging of point range " + std::to_string(r) + " refused (test hook)"};
                    stage(r, lo[r], lo[r + 1] - lo[r]);
                    HIPCK(hipEventRecord(cx.range_ready[r], cx.stream));
                    std::lock_guard<std::mutex> lk(mu_);
                    recorded_ = r + 1;
                    cv_.notify_all();
                }
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu_);
                error_ = std::current_exception();
                recorded_ = K_;
                cv_.notify_all();
            }
        });
    }
    ~RangeStager() {
        if (thread_.joinable()) thread_.join();
    }
    void wait(uint32_t r) {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return recorded_ > r; });
        if (error_) std::rethrow_exception(error_);
    }
    void join() {
        if (thread_.joinable()) thread_.join();
        if (error_) std::rethrow_exception(error_);
    }

private:
    const uint32_t K_;
    std::thread thread_;
    std::mutex mu_;
    std::condition_variable cv_;
    uint32_t recorded_ = 0;
    std::exception_ptr error_;
};

// An MSM in K point ranges whose buckets are shared: each window-group
// pipeline sorts and accumulates range r (the accumulation continues the
// bucket sums of the earlier ranges, their entry counts summed alongside)
// and reduces the buckets once, after the last range.  full gives the
// window and segment length of the whole MSM, a range takes the tile of its
// own size (the whole size's tile would leave the SIMDs short of waves).
// stage(r, lo, m) runs before range r's pipelines are queued and leaves the
// range's inputs (m points from point lo) written on cx.stream.  Range bounds
// fall on multiples of 64 points, so every range's scalars stay 64-byte
// aligned whatever the scalar width.
template <class Stage>
void ranged_msm(Ctx& cx, const JobSpec& full, size_t nbytes, uint32_t K, Stage&& stage, blst_p1* ret) {
    const size_t n = full.n;
    const uint8_t* dp = static_cast<const uint8_t*>(full.points);
    const uint8_t* ds = static_cast<const uint8_t*>(full.scalars);
    const uint32_t G = pipeline_groups(full);
    const uint32_t nwin = 1 + full.nbits / full.c;
    while (cx.pipe.size() < G) {
        cx.pipe.emplace_back(new Workspace());
        cx.pipe.back()->init();
    }
    while (cx.range_ready.size() < K) {
        hipEvent_t e;
        HIPCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        cx.range_ready.push_back(e);
    }
    std::vector<size_t> lo(K + 1);
    for (uint32_t k = 0; k < K; k++) lo[k] = (n * k / K) & ~(size_t)63;
    lo[K] = n;
    // the first range's copies start while the layouts are built
    RangeStager staging(cx, lo, stage);
    // layouts of every (group, range), buffers and tables first (allocation
    // may synchronise); each range's tables in a region of their own
    std::vector<Layout> Ls(G * K);
    std::vector<DevLayout> Ds(G * K);
    std::vector<hostec::xyzz6*> outs(G);
    std::vector<uint32_t> nsums(G);
    for (uint32_t g = 0; g < G; g++) {
        Workspace& w = *cx.pipe[g];
        size_t stride = 0;
        for (uint32_t r = 0; r < K; r++) {
            JobSpec s = full;
            s.points = dp + lo[r] * sizeof(blst_p1_affine);
            s.scalars = ds + lo[r] * nbytes;
            s.n = (uint32_t)(lo[r + 1] - lo[r]);
            s.T = make_spec(s.points, s.scalars, s.n, full.nbits, full.c, false).T;
            s.win0 = group_win0(g, G, nwin);
            s.nwin = group_win_end(g, G, nwin) - s.win0;
            Layout& L = Ls[g * K + r];
            L = make_layout(std::vector<JobSpec>(1, s));
            ensure_workspace(w, L);
            stride = std::max(stride, layout_table_bytes(L));
        }
        w.tables.ensure(K * stride);
        (void)w.tables_stage(K * stride);
        w.cnt_sum.ensure(((size_t)Ls[g * K].nkeys + 1) * 4);
        for (uint32_t r = 0; r < K; r++) Ds[g * K + r] = upload_layout(w, Ls[g * K + r], r * stride);
        nsums[g] = (uint32_t)Ls[g * K].groups.size();
        outs[g] = static_cast<hostec::xyzz6*>(w.host_buf(nsums[g] * sizeof(hostec::xyzz6)));
    }
    for (uint32_t r = 0; r < K; r++) {
        staging.wait(r);
        for (uint32_t g = 0; g < G; g++) {
            Workspace& w = *cx.pipe[g];
            HIPCK(hipStreamWaitEvent(w.stream, cx.range_ready[r], 0));
            RangeRole role;
            role.prev = r ? w.cnt_sum.as<const uint32_t>() : nullptr;
            role.cnt_sum = w.cnt_sum.as<uint32_t>();
            role.reduce = r + 1 == K;
            enqueue_batch(w, Ls[g * K + r], Ds[g * K + r], true, outs[g], g == 0 && r + 1 == K, role);
        }
    }
    staging.join();
    fold_groups(cx, G, nwin, full.c, outs, nsums, ret);
}

// An MSM over contiguous host buffers (points, and scalars of nbits bits or,
// with fr, Montgomery-form blst_fr values converted on the device).  Its
// inputs cross PCIe (134 MB at 2^20 points, about as long as the MSM itself),
// so from kHostSplitMin points they are copied in ranges of about
// kHostRangePoints points, range r + 1 staged and copied while the pipelines
// work on range r (ranged_msm).
constexpr size_t kHostSplitMin = (size_t)1 << 19;
constexpr size_t kHostRangePoints = (size_t)1 << 18;   // 2 to 16 ranges

uint32_t point_ranges(size_t n, size_t min_n, size_t range_points) {
    if (n < min_n) return 1;
    return (uint32_t)std::min<size_t>(16, std::max<size_t>(2, n / range_points));
}

constexpr size_t kDeviceSplitMin = (size_t)1 << 21;
constexpr size_t kDeviceRangePoints = (size_t)1 << 19;

void host_buffers_msm(Ctx& cx, const uint8_t* hp, const uint8_t* hs, size_t n, size_t nbits, size_t window,
                      bool fr, bool fold, size_t bit0, blst_p1* ret) {
    const size_t nbytes = fr ? 32 : (nbits + 7) / 8;
    cx.pts.ensure(n * sizeof(blst_p1_affine));
    cx.scl.ensure(n * nbytes + 16);
    uint8_t* dp = static_cast<uint8_t*>(cx.pts.p);
    uint8_t* ds = static_cast<uint8_t*>(cx.scl.p);
    auto stage = [&](uint32_t, size_t lo, size_t m) {
        cx.stager.upload(ds + lo * nbytes, hs + lo * nbytes, m * nbytes, cx.stream);
        cx.stager.upload(dp + lo * sizeof(blst_p1_affine), hp + lo * sizeof(blst_p1_affine),
                         m * sizeof(blst_p1_affine), cx.stream);
        if (fr) launch_k_fr_to_scalar(cx.stream, reinterpret_cast<uint32_t*>(ds + lo * nbytes), (uint32_t)m);
        HIPCK(hipGetLastError());
    };
    JobSpec js = make_spec(dp, ds, n, nbits, window, false);
    const uint32_t K = fold ? point_ranges(n, kHostSplitMin, kHostRangePoints) : 1;
    if (K > 1) {
        ranged_msm(cx, js, nbytes, K, stage, ret);
        return;
    }
    stage(0, 0, n);
    js.bit_base = (uint32_t)bit0;
    js.tile = !fold;
    device_msm(cx, js, ret);
}

// Resolves the reference's descriptor convention (a NULL entry after the head
// means: continue contiguously from the previous element,
// blst/scalar/curve/multi_scalar.cpp:838-841) into one contiguous copy.
template <class T>
const uint8_t* resolve(const T* const items[], size_t n, size_t stride, std::vector<uint8_t>& stage) {
    const uint8_t* head = reinterpret_cast<const uint8_t*>(items[0]);
    bool contiguous = true, ended = false;
    const uint8_t* prev = head;
    for (size_t i = 1; i < n && contiguous; i++) {
        if (!items[i]) break;   // the rest continues contiguously: {base, NULL} is one check
        const uint8_t* cur = reinterpret_cast<const uint8_t*>(items[i]);
        if (cur != prev + stride) contiguous = false;
        prev = cur;
    }
    if (contiguous) return head;
    stage.resize(n * stride);
    ended = false;
    prev = head;
    std::memcpy(stage.data(), head, stride);
    for (size_t i = 1; i < n; i++) {
        const uint8_t* cur;
        if (!ended && items[i]) cur = reinterpret_cast<const uint8_t*>(items[i]);
        else { ended = true; cur = prev + stride; }
        std::memcpy(stage.data() + i * stride, cur, stride);
        prev = cur;
    }
    return stage.data();
}

int host_msm(blst_p1* ret, const blst_p1_affine* const points[], size_t npoints,
             const blst_byte* const scalars[], size_t nbits, size_t window, bool fold,
             size_t bit0) {
    auto t0 = std::chrono::steady_clock::now();
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        const size_t nbytes = (nbits + 7) / 8;
        std::vector<uint8_t> sp, ss;
        const uint8_t* hp = resolve(points, npoints, sizeof(blst_p1_affine), sp);
        const uint8_t* hs = resolve(scalars, npoints, nbytes, ss);
        host_buffers_msm(cx, hp, hs, npoints, nbits, window, false, fold, bit0, ret);
    } catch (...) {
        g_err = current_error();
        std::memset(ret, 0, sizeof(*ret));
        return -1;
    }
    g_phase[6] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    g_err.clear();
    return 0;
}

// A G2 MSM (one pipeline over Fp2 points) on inputs already on the device;
// the host folds the windows over Fp2.
void device_msm_g2(Ctx& cx, const JobSpec& js, blst_p2* ret) {
    std::vector<JobSpec> v(1, js);
    v[0].g2 = true;
    Layout L = make_layout(v);
    ensure_workspace(cx.ws, L);
    const DevLayout D = upload_layout(cx.ws, L);
    std::vector<hostec2::xyzz12> wins(L.groups.size());
    enqueue_batch(cx.ws, L, D, true, wins.data());
    HIPCK(hipStreamSynchronize(cx.stream));
    record_phases(cx.ws);
    hostec2::xyzz12 r;
    if (!js.tile) hostec2::fold_windows(r, wins.data(), (int)wins.size(), (int)js.c);
    else r = wins[0];
    hostec2::jac12 j;
    hostec2::to_jac(j, r);
    static_assert(sizeof(hostec2::jac12) == sizeof(blst_p2), "blst_p2 layout");
    std::memcpy(ret, &j, sizeof(j));
}

// A G2 MSM in K point ranges sharing one set of buckets, as ranged_msm does
// for G1: one window group, on a pipeline stream of its own so that the
// copies of the next range (queued on cx.stream by stage) run beside it.
template <class Stage>
void ranged_msm_g2(Ctx& cx, const JobSpec& full, size_t nbytes, uint32_t K, Stage&& stage, blst_p2* ret) {
    const size_t n = full.n;
    const uint8_t* dp = static_cast<const uint8_t*>(full.points);
    const uint8_t* ds = static_cast<const uint8_t*>(full.scalars);
    if (cx.pipe.empty()) {
        cx.pipe.emplace_back(new Workspace());
        cx.pipe.back()->init();
    }
    Workspace& w = *cx.pipe[0];
    while (cx.range_ready.size() < K) {
        hipEvent_t e;
        HIPCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        cx.range_ready.push_back(e);
    }
    std::vector<size_t> lo(K + 1);
    for (uint32_t k = 0; k < K; k++) lo[k] = (n * k / K) & ~(size_t)63;
    lo[K] = n;
    RangeStager staging(cx, lo, stage);   // copies start while the layouts are built
    std::vector<Layout> Ls(K);
    std::vector<DevLayout> Ds(K);
    size_t stride = 0;
    for (uint32_t r = 0; r < K; r++) {
        JobSpec s = full;
        s.points = dp + lo[r] * sizeof(blst_p2_affine);
        s.scalars = ds + lo[r] * nbytes;
        s.n = (uint32_t)(lo[r + 1] - lo[r]);
        s.T = make_spec(s.points, s.scalars, s.n, full.nbits, full.c, false).T;
        s.g2 = true;
        Ls[r] = make_layout(std::vector<JobSpec>(1, s));
        ensure_workspace(w, Ls[r]);
        stride = std::max(stride, layout_table_bytes(Ls[r]));
    }
    w.tables.ensure(K * stride);
    (void)w.tables_stage(K * stride);
    w.cnt_sum.ensure(((size_t)Ls[0].nkeys + 1) * 4);
    for (uint32_t r = 0; r < K; r++) Ds[r] = upload_layout(w, Ls[r], r * stride);
    const uint32_t nsum = (uint32_t)Ls[0].groups.size();
    auto* out = static_cast<hostec2::xyzz12*>(w.host_buf(nsum * sizeof(hostec2::xyzz12)));
    for (uint32_t r = 0; r < K; r++) {
        staging.wait(r);
        HIPCK(hipStreamWaitEvent(w.stream, cx.range_ready[r], 0));
        RangeRole role;
        role.prev = r ? w.cnt_sum.as<const uint32_t>() : nullptr;
        role.cnt_sum = w.cnt_sum.as<uint32_t>();
        role.reduce = r + 1 == K;
        enqueue_batch(w, Ls[r], Ds[r], true, out, r + 1 == K, role);
    }
    staging.join();
    HIPCK(hipStreamSynchronize(w.stream));
    record_phases(w);
    hostec2::xyzz12 res;
    hostec2::fold_windows(res, out, (int)nsum, (int)full.c);
    hostec2::jac12 j;
    hostec2::to_jac(j, res);
    std::memcpy(ret, &j, sizeof(j));
}

// A G2 MSM over contiguous host buffers (points, and scalars of nbits bits
// or, with fr, Montgomery-form blst_fr values converted on the device): from
// host buffers the points cross PCIe in ranges, as for G1.
void host_buffers_msm_g2(Ctx& cx, const uint8_t* hp, const uint8_t* hs, size_t npoints, size_t nbits, size_t window,
                         bool fr, bool fold, size_t bit0, blst_p2* ret) {
    const size_t nbytes = fr ? 32 : (nbits + 7) / 8;
    cx.pts.ensure(npoints * sizeof(blst_p2_affine));
    cx.scl.ensure(npoints * nbytes + 16);
    uint8_t* dp = static_cast<uint8_t*>(cx.pts.p);
    uint8_t* ds = static_cast<uint8_t*>(cx.scl.p);
    auto stage = [&](uint32_t, size_t lo, size_t m) {
        cx.stager.upload(ds + lo * nbytes, hs + lo * nbytes, m * nbytes, cx.stream);
        cx.stager.upload(dp + lo * sizeof(blst_p2_affine), hp + lo * sizeof(blst_p2_affine),
                         m * sizeof(blst_p2_affine), cx.stream);
        if (fr) launch_k_fr_to_scalar(cx.stream, reinterpret_cast<uint32_t*>(ds + lo * nbytes), (uint32_t)m);
        HIPCK(hipGetLastError());
    };
    JobSpec js = make_spec(dp, ds, npoints, nbits, window, false);
    const uint32_t K = fold ? point_ranges(npoints, kHostSplitMin, kHostRangePoints) : 1;
    if (K > 1) {
        ranged_msm_g2(cx, js, nbytes, K, stage, ret);
    } else {
        stage(0, 0, npoints);
        js.bit_base = (uint32_t)bit0;
        js.tile = !fold;
        device_msm_g2(cx, js, ret);
    }
}

int host_msm_g2(blst_p2* ret, const blst_p2_affine* const points[], size_t npoints,
                const blst_byte* const scalars[], size_t nbits, size_t window, bool fold, size_t bit0) {
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        const size_t nbytes = (nbits + 7) / 8;
        std::vector<uint8_t> sp, ss;
        const uint8_t* hp = resolve(points, npoints, sizeof(blst_p2_affine), sp);
        const uint8_t* hs = resolve(scalars, npoints, nbytes, ss);
        host_buffers_msm_g2(cx, hp, hs, npoints, nbits, window, false, fold, bit0, ret);
    } catch (...) {
        g_err = current_error();
        std::memset(ret, 0, sizeof(*ret));
        return -1;
    }
    g_err.clear();
    return 0;
}

bool heads_ok(const void* const* xs, const void* const* ys, size_t n) {
    return n == 0 || (xs != nullptr && ys != nullptr && xs[0] != nullptr && ys[0] != nullptr);
}

// A batch captured once as a HIP graph: fixed job shapes and fixed device
// input pointers; each run replays the whole launch sequence.
struct PlanImpl {
    int device = 0;
    Workspace ws;
    Layout L;
    DevLayout D;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hostec::xyzz6* host_res = nullptr;   // pinned, one per nonempty job
    std::vector<size_t> idx;             // batch position of each nonempty job
    size_t njobs = 0;
    ~PlanImpl() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (host_res) (void)hipHostFree(host_res);
    }
};

bool specs_from_jobs(const blst_hip_msm_job* jobs, size_t njobs, std::vector<JobSpec>& specs,
                     std::vector<size_t>& idx) {
    for (size_t k = 0; k < njobs; k++) {
        const blst_hip_msm_job& j = jobs[k];
        if (j.npoints == 0 || j.nbits == 0) continue;   // infinity, as the reference
        if (j.nbits > 256 || j.window > 24 || !j.points || !j.scalars || ((uintptr_t)j.points % 16) != 0 ||
            j.npoints >= 0x7fffffffULL) {
            g_err = "invalid job " + std::to_string(k);
            return false;
        }
        JobSpec js = make_spec(j.points, j.scalars, j.npoints, j.nbits, j.window, true);
        specs.push_back(js);
        idx.push_back(k);
    }
    return true;
}

// Lanes of the Jacobian -> affine kernel: four waves per CU, each lane
// inverting once for every n / 65536 points.
constexpr size_t kAffineLanes = 65536;
// lanes of the XYZZ -> affine conversion of table rows and comb tables
// (131,072 against 65,536: comb table of 4096 points at wbits 13 built in 190
// against 205 ms; 262,144 no better)
constexpr size_t kTableAffineLanes = 131072;
// points per host round trip of blst_p1s_to_affine (2.4 GB in, 1.6 GB out)
constexpr size_t kAffineChunk = (size_t)1 << 24;

template <class F = fp>
void enqueue_to_affine(hipStream_t stream, const void* d_in, void* d_out, size_t n) {
    const uint32_t T = (uint32_t)std::min(n, kAffineLanes);
    launch_k_jac_to_affine(stream, static_cast<const uint4*>(d_in), static_cast<aff_t<F>*>(d_out), (uint32_t)n, T);
    HIPCK(hipGetLastError());
}

// ---------------------------------------------------------------- fixed-base tables

constexpr size_t kWbitsRowEntries = (size_t)1 << 22;   // table entries per row-building launch
constexpr size_t kWbitsSumLanes = 65536;
constexpr size_t kWbitsListBytes = (size_t)256 << 20;   // gathered entries per host round trip

bool valid_wbits(size_t wbits) { return wbits >= 2 && wbits <= 14; }

// The groups of the fixed-base tables: point layouts, the device field and
// the host's XYZZ arithmetic for the window fold.
struct G1Ops {
    using Aff = blst_p1_affine;
    using Jac = blst_p1;
    using F = fp;
    using HX = hostec::xyzz6;
    static void fold(HX& r, const HX* w, int n, int c) { hostec::fold_windows(r, w, n, c); }
    static void addp(HX& r, const HX& a, const HX& b) { hostec::addp(r, a, b); }
    static void finish(Jac* ret, const HX& x) {
        hostec::jac6 j;
        hostec::to_jac(j, x);
        std::memcpy(ret, &j, sizeof(j));
    }
    // y = -y of a staged affine point (0 stays 0)
    static void neg_y(uint8_t* pt) {
        hostec::fp6 zero{}, y;
        std::memcpy(&y, pt + 48, 48);
        hostec::sub(y, zero, y);
        std::memcpy(pt + 48, &y, 48);
    }
};

struct G2Ops {
    using Aff = blst_p2_affine;
    using Jac = blst_p2;
    using F = fp2;
    using HX = hostec2::xyzz12;
    static void fold(HX& r, const HX* w, int n, int c) { hostec2::fold_windows(r, w, n, c); }
    static void addp(HX& r, const HX& a, const HX& b) { hostec2::addp(r, a, b); }
    static void finish(Jac* ret, const HX& x) {
        hostec2::jac12 j;
        hostec2::to_jac(j, x);
        std::memcpy(ret, &j, sizeof(j));
    }
    static void neg_y(uint8_t* pt) {
        hostec::fp6 zero{}, y;
        for (int c = 0; c < 2; c++) {
            std::memcpy(&y, pt + 96 + 48 * c, 48);
            hostec::sub(y, zero, y);
            std::memcpy(pt + 96 + 48 * c, &y, 48);
        }
    }
};

// rows of n points (device, affine) into table (device, n * 2^(wbits-1)
// affine entries)
template <class G = G1Ops>
void enqueue_wbits_table(Ctx& cx, const void* d_pts, size_t n, size_t wbits, void* d_table) {
    using X = xyzz_t<typename G::F>;
    using A = aff_t<typename G::F>;
    const size_t m = (size_t)1 << (wbits - 1);
    const uint32_t L = (uint32_t)std::min<size_t>(m, 64);
    const size_t R = std::max<size_t>(1, kWbitsRowEntries / m);
    cx.wb_rows.ensure(std::min(n, R) * m * sizeof(X));
    for (size_t row0 = 0; row0 < n; row0 += R) {
        const size_t nr = std::min(R, n - row0);
        const size_t count = nr * m;
        launch_k_wbits_rows(cx.stream, static_cast<const uint4*>(d_pts), (uint32_t)row0, (uint32_t)nr, (uint32_t)m, L,
                            cx.wb_rows.as<X>());
        const uint32_t Q = (uint32_t)std::min<size_t>(64, std::max<size_t>(1, count / kTableAffineLanes));
        launch_k_to_affine(dim3(div_up(div_up(count, Q), 256)), dim3(256), cx.stream, cx.wb_rows.as<const X>(),
                           static_cast<A*>(d_table) + row0 * m, (uint32_t)count, Q);
        HIPCK(hipGetLastError());
    }
}

// window sums S_w of n (window, point) entries, from a device table and
// device scalars or from a gathered list (window-major, D * n entries),
// added into W (D windows)
template <class G = G1Ops>
void wbits_window_sums(Ctx& cx, const void* d_table, const void* d_scl, size_t nbits, const void* d_list, size_t n,
                       size_t wbits, std::vector<typename G::HX>& W) {
    using X = xyzz_t<typename G::F>;
    const size_t D = 1 + nbits / wbits;
    const size_t nbytes = (nbits + 7) / 8;
    const size_t C = std::min(n, std::max<size_t>(4, (n * D + kWbitsSumLanes - 1) / kWbitsSumLanes));
    const size_t nch = (n + C - 1) / C;
    cx.wb_part.ensure(D * nch * sizeof(X));
    cx.wb_W.ensure(D * sizeof(X));
    const uint32_t vec = (nbytes == 32 && ((uintptr_t)d_scl % 16) == 0) ? 1u : 0u;
    launch_k_wbits_sum(cx.stream, static_cast<const uint4*>(d_table), static_cast<const uint8_t*>(d_scl),
                       (uint32_t)nbytes, (uint32_t)nbits, vec, static_cast<const uint4*>(d_list), (uint32_t)n,
                       (uint32_t)wbits, (uint32_t)D, (uint32_t)C, (uint32_t)nch, cx.wb_part.as<X>());
    launch_k_wbits_tree(cx.stream, cx.wb_part.as<const X>(), (uint32_t)nch, (uint32_t)D, cx.wb_W.as<X>());
    HIPCK(hipGetLastError());
    static_assert(sizeof(typename G::HX) == sizeof(X), "host and device XYZZ layouts");
    std::vector<typename G::HX> part(D);
    HIPCK(hipMemcpyAsync(part.data(), cx.wb_W.p, D * sizeof(X), hipMemcpyDeviceToHost, cx.stream));
    HIPCK(hipStreamSynchronize(cx.stream));
    if (W.empty()) W.assign(D, typename G::HX{});
    for (size_t w = 0; w < D; w++) G::addp(W[w], W[w], part[w]);
}

// Booth digit of the window at bit0 (<= nbits) of width c, from the scalar's
// bytes (the device's booth_digit_at)
int32_t host_booth_digit(const uint8_t* s, uint32_t bit0, uint32_t c, uint32_t nbits) {
    auto bits = [&](uint32_t pos, uint32_t len) -> uint32_t {
        const uint32_t b0 = pos >> 3, b1 = (pos + len - 1) >> 3;
        uint64_t v = 0;
        for (uint32_t b = b0; b <= b1; b++) v |= (uint64_t)s[b] << (8 * (b - b0));
        return (uint32_t)(v >> (pos & 7)) & ((1u << len) - 1u);
    };
    const uint32_t e = std::min(nbits - bit0 + 1, c);
    const uint32_t end = std::min(bit0 + c, nbits);
    const uint32_t wval = bit0 == 0 ? bits(0, end) << 1 : bits(bit0 - 1, end - bit0 + 1);
    const uint32_t sign = (wval >> e) & 1u;
    const int32_t val = (int32_t)((wval + 1) >> 1);
    return sign ? val - (int32_t)(1u << e) : val;
}

void* wbits_host_stage(Ctx& cx, size_t bytes) {
    if (bytes > cx.wb_host_cap) {
        if (cx.wb_host) HIPCK(hipHostFree(cx.wb_host));
        cx.wb_host = nullptr;
        cx.wb_host_cap = 0;
        HIPCK(hipHostMalloc(&cx.wb_host, bytes, hipHostMallocDefault));
        cx.wb_host_cap = bytes;
    }
    return cx.wb_host;
}

// The host-table MSM: the host computes digits and gathers the signed
// entries of a run of points for every window into page-locked memory (the
// table stays where the caller keeps it), the device sums them.
template <class G = G1Ops>
void host_mult_wbits(Ctx& cx, typename G::Jac* ret, const typename G::Aff* table, size_t wbits, size_t n,
                     const uint8_t* scalars, size_t nbits) {
    using Aff = typename G::Aff;
    const size_t D = 1 + nbits / wbits, nbytes = (nbits + 7) / 8, m = (size_t)1 << (wbits - 1);
    const size_t run = std::max<size_t>(1, std::min(n, kWbitsListBytes / (D * sizeof(Aff))));
    uint8_t* stage = static_cast<uint8_t*>(wbits_host_stage(cx, run * D * sizeof(Aff)));
    cx.wb_list.ensure(run * D * sizeof(Aff));
    std::vector<typename G::HX> W;
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (size_t i0 = 0; i0 < n; i0 += run) {
        const size_t cnt = std::min(run, n - i0);
        auto gather = [&](size_t a, size_t b) {
            for (size_t i = a; i < b; i++) {
                const uint8_t* sc = scalars + (i0 + i) * nbytes;
                const Aff* row = table + (i0 + i) * m;
                for (size_t w = 0; w < D; w++) {
                    uint8_t* dst = stage + (w * cnt + i) * sizeof(Aff);
                    const int32_t d = host_booth_digit(sc, (uint32_t)(w * wbits), (uint32_t)wbits, (uint32_t)nbits);
                    if (d == 0) {
                        std::memset(dst, 0, sizeof(Aff));
                        continue;
                    }
                    std::memcpy(dst, &row[(d < 0 ? -d : d) - 1], sizeof(Aff));
                    if (d < 0) G::neg_y(dst);
                }
            }
        };
        const size_t per = (cnt + nthreads - 1) / nthreads;
        if (cnt * D < 4096 || nthreads == 1) {
            gather(0, cnt);
        } else {
            std::vector<std::thread> th;
            for (unsigned t = 0; t < nthreads; t++) {
                const size_t a = t * per, b = std::min(cnt, a + per);
                if (a < b) th.emplace_back(gather, a, b);
            }
            for (auto& t : th) t.join();
        }
        HIPCK(hipMemcpyAsync(cx.wb_list.p, stage, cnt * D * sizeof(Aff), hipMemcpyHostToDevice, cx.stream));
        wbits_window_sums<G>(cx, nullptr, nullptr, nbits, cx.wb_list.p, cnt, wbits, W);
    }
    typename G::HX r;
    G::fold(r, W.data(), (int)D, (int)wbits);
    G::finish(ret, r);
}

// The fixed-base entry points of either group (blst/c_api.cpp:1310-1369 for
// G1, 1829-1888 for G2): the reference's checks in its order, then the
// device table build or the host-table MSM.
template <class G>
size_t wbits_precompute_bytes(size_t wbits, size_t npoints) {
    if (!valid_wbits(wbits)) return 0;
    size_t bytes;
    if (__builtin_mul_overflow(npoints, sizeof(typename G::Aff), &bytes)) return 0;
    if (bytes > (SIZE_MAX >> (wbits - 1))) return 0;
    return bytes << (wbits - 1);
}

template <class G>
void wbits_precompute_entry(typename G::Aff table[], size_t wbits, const typename G::Aff* const points[],
                            size_t npoints) {
    using Aff = typename G::Aff;
    if (!valid_wbits(wbits) || (npoints != 0 && table == nullptr) ||
        (npoints != 0 && (points == nullptr || points[0] == nullptr))) {
        g_err = "invalid arguments";
        return;
    }
    if (npoints == 0) { g_err.clear(); return; }
    if (wbits_precompute_bytes<G>(wbits, npoints) == 0 || npoints >= 0x7fffffffULL) {
        g_err = "table size out of range";
        return;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        std::vector<uint8_t> stage;
        const uint8_t* hp = resolve(points, npoints, sizeof(Aff), stage);
        const size_t m = (size_t)1 << (wbits - 1);
        cx.wb_pts.ensure(npoints * sizeof(Aff));
        // every point is on the device before any row is written back: a
        // packed input may occupy the table's tail
        cx.stager.upload(cx.wb_pts.p, hp, npoints * sizeof(Aff), cx.stream);
        const size_t R = std::max<size_t>(1, kWbitsRowEntries / m);
        cx.wb_table.ensure(std::min(npoints, R) * m * sizeof(Aff));
        for (size_t row0 = 0; row0 < npoints; row0 += R) {
            const size_t nr = std::min(R, npoints - row0);
            const uint8_t* rows = static_cast<const uint8_t*>(cx.wb_pts.p) + row0 * sizeof(Aff);
            enqueue_wbits_table<G>(cx, rows, nr, wbits, cx.wb_table.p);
            cx.stager.download(table + row0 * m, cx.wb_table.p, nr * m * sizeof(Aff), cx.stream);
        }
    } catch (...) {
        g_err = current_error();
        return;
    }
    g_err.clear();
}

template <class G>
void mult_wbits_entry(typename G::Jac* ret, const typename G::Aff table[], size_t wbits, size_t npoints,
                      const blst_byte* const scalars[], size_t nbits, blst_limb_t* scratch) {
    if (npoints == 0 || nbits == 0) { std::memset(ret, 0, sizeof(*ret)); g_err.clear(); return; }
    if (nbits > 256 || !valid_wbits(wbits) || table == nullptr || scalars == nullptr || scalars[0] == nullptr ||
        scratch == nullptr) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "invalid arguments";
        return;
    }
    if (wbits_precompute_bytes<G>(wbits, npoints) == 0 || npoints >= 0x7fffffffULL) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "table size out of range";
        return;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        std::vector<uint8_t> stage;
        const uint8_t* hs = resolve(scalars, npoints, (nbits + 7) / 8, stage);
        host_mult_wbits<G>(cx, ret, table, wbits, npoints, hs, nbits);
    } catch (...) {
        g_err = current_error();
        std::memset(ret, 0, sizeof(*ret));
        return;
    }
    g_err.clear();
}

// blst_p2s_to_affine (blst/c_api.cpp:1806-1817): chunks of the Jacobian
// points up, the batch normalisation, the affine points down
void p2s_to_affine(blst_p2_affine dst[], const blst_p2* const points[], size_t npoints) {
    if (npoints == 0) { g_err.clear(); return; }
    if (dst == nullptr) { g_err = "dst is NULL"; return; }
    if (points == nullptr || points[0] == nullptr) {
        std::memset(dst, 0, npoints * sizeof(blst_p2_affine));
        g_err.clear();
        return;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        std::vector<uint8_t> stage;
        const uint8_t* hp = resolve(points, npoints, sizeof(blst_p2), stage);
        const size_t chunk = std::min(npoints, kAffineChunk / 2);
        cx.jac_in.ensure(chunk * sizeof(blst_p2));
        cx.aff_out.ensure(chunk * sizeof(blst_p2_affine));
        for (size_t b = 0; b < npoints; b += chunk) {
            const size_t m = std::min(chunk, npoints - b);
            cx.stager.upload(cx.jac_in.p, hp + b * sizeof(blst_p2), m * sizeof(blst_p2), cx.stream);
            enqueue_to_affine<fp2>(cx.stream, cx.jac_in.p, cx.aff_out.p, m);
            cx.stager.download(dst + b, cx.aff_out.p, m * sizeof(blst_p2_affine), cx.stream);
        }
    } catch (...) {
        g_err = current_error();
        std::memset(dst, 0, npoints * sizeof(blst_p2_affine));
        return;
    }
    g_err.clear();
}

// ---------------------------------------------------------------- comb tables

// bytes of a comb table: D windows of n rows of 2^(wbits-1) affine entries;
// 0 for invalid parameters or on overflow
size_t comb_table_bytes(size_t npoints, size_t wbits, size_t nbits) {
    if (!valid_wbits(wbits) || nbits == 0 || nbits > 256 || npoints == 0) return 0;
    const size_t D = 1 + nbits / wbits;
    size_t b;
    if (__builtin_mul_overflow(npoints, D * sizeof(blst_p1_affine), &b)) return 0;
    if (b > (SIZE_MAX >> (wbits - 1))) return 0;
    return b << (wbits - 1);
}

// ---------------------------------------------------------------- no-allocation MSM
//
// blst_hip_msm_device_noalloc is blst_hip_msm_device for at most
// kNoAllocMaxPoints points with every device buffer of its batch carved out
// of `scratch`, a fixed array of this translation unit: the call allocates no
// device or page-locked memory.  Each GPU the library runs on holds its own
// copy of the array (kNoAllocScratch bytes) until the process ends.  In host
// code `scratch` names a host-side stand-in, not a device address: the device
// address of each GPU's copy comes from hipGetSymbolAddress, once per device
// with that device current, and only that address is handed to kernels.
constexpr size_t kNoAllocMaxPoints = (size_t)1 << 16;
constexpr size_t kNoAllocScratch = (size_t)96 << 20;   // a size class whose tuned tile needs more takes a wider tile (noalloc_msm)
// host staging of the layout tables, then the window sums
constexpr size_t kNoAllocHostBytes = (size_t)1 << 20;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratch];

struct NoAllocDevice {
    std::mutex mu;               // one call at a time on this device
    uint8_t* base = nullptr;     // device address of this GPU's scratch
    Workspace ws;                // buffers attached inside [base, base + kNoAllocScratch)
    std::vector<uint8_t> host;   // fixed host staging, sized at set-up
};

// Set-up on the first call on a device, with that device current: the
// address of its copy of scratch and the stream of its calls.  Kept for the
// life of the process (the runtime may be gone when static objects die).
NoAllocDevice& noalloc_device() {
    static std::mutex mu;
    static auto* devices = new std::map<int, NoAllocDevice*>();
    int dev = 0;
    HIPCK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = devices->find(dev);
    if (it != devices->end()) return *it->second;
    std::unique_ptr<NoAllocDevice> d(new NoAllocDevice());
    void* ptr = nullptr;
    HIPCK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr || ((uintptr_t)ptr & 255)) throw HipFail{"scratch array has no usable device address"};
    d->base = static_cast<uint8_t*>(ptr);
    // the device context's stream (created with the context, once per
    // device): a stream of its own would take one of the few hardware queues
    d->ws.stream = context().stream;
    d->host.resize(kNoAllocHostBytes);
    d->ws.host_fixed = d->host.data();
    d->ws.host_fixed_cap = kNoAllocHostBytes / 2;
    NoAllocDevice* raw = d.release();
    devices->emplace(dev, raw);
    return *raw;
}

// The MSM of blst_hip_msm_device (tuned window, one pipeline: every size up
// to kNoAllocMaxPoints runs as one window group) in this device's scratch.
// Where the tuned tile's partial-sum slots (two per tile) would not fit the
// fixed region, the tile doubles until they do: the tile sets only how the
// sorted entries are cut among lanes, never the result.
void noalloc_msm(const void* device_points, const void* device_scalars, size_t npoints, size_t nbits,
                 blst_p1* ret) {
    NoAllocDevice& d = noalloc_device();
    std::lock_guard<std::mutex> lk(d.mu);
    JobSpec js = make_spec(device_points, device_scalars, npoints, nbits, 0, false);
    Layout L = make_layout(std::vector<JobSpec>(1, js));
    while (carved_bytes(L) > kNoAllocScratch && js.T < 4096) {
        js.T *= 2;
        L = make_layout(std::vector<JobSpec>(1, js));
    }
    carve_workspace(d.ws, L, d.base, kNoAllocScratch);
    const size_t nsums = L.groups.size();
    if (nsums * sizeof(hostec::xyzz6) > kNoAllocHostBytes / 2) throw HipFail{"too many windows"};
    hostec::xyzz6* sums = reinterpret_cast<hostec::xyzz6*>(d.host.data() + kNoAllocHostBytes / 2);
    const DevLayout D = upload_layout(d.ws, L);
    enqueue_batch(d.ws, L, D, true, sums, false);
    HIPCK(hipStreamSynchronize(d.ws.stream));
    hostec::xyzz6 r;
    hostec::fold_windows(r, sums, (int)nsums, (int)js.c);
    hostec::jac6 j;
    hostec::to_jac(j, r);
    std::memcpy(ret, &j, sizeof(j));
}

}  // namespace host
}  // namespace bls

using namespace bls::host;

struct blst_hip_batch_plan {
    PlanImpl impl;
};

// ---------------------------------------------------------------- Fr scalars
//
// Scratch sizes of the reference's blst_fr entry points, which callers
// allocate from: 32 bytes per canonical scalar plus the Pippenger buckets
// (blst/c_api.cpp:272-299), and for blst_p1s_mult_fr the largest of that and
// the batch-affine plan's workspace (c_api.cpp:593-627, 705-807,
// blst/scalar/curve/multi_scalar_batch_affine.cpp:16-36, 149-168).  Overflow
// yields 0.

namespace {

// The element sizes of one group's scratch layouts: the Pippenger bucket
// (Pointxyzz), and the batch-affine workspace's affine bucket (x, y, two
// u32), projective collision bucket (Jacobian point + u32, 8-byte aligned)
// and field element; the largest size the reference's blst_fr planner sends
// to batch-affine (multi_scalar_batch_affine.cpp:16-36,
// c_api.cpp:700-719).
struct FrLayout {
    size_t xyzz, affine_bucket, projective_bucket, field, batch_affine_max;
};
constexpr FrLayout kFrG1 = {192, 104, 152, 48, (size_t)1 << 18};
constexpr FrLayout kFrG2 = {384, 200, 296, 96, (size_t)1 << 14};

size_t pippenger_fr_scratch(size_t n, const FrLayout& G = kFrG1) {
    if (n == 0) return 0;
    size_t scalar_bytes;
    if (__builtin_mul_overflow(n, (size_t)32, &scalar_bytes)) return 0;
    if (n == 1) return scalar_bytes;
    const size_t w = ref_window_size(n);
    if (w == 0 || w - 1 >= 8 * sizeof(size_t)) return 0;
    size_t buckets, total;
    if (__builtin_mul_overflow((size_t)1 << (w - 1), G.xyzz, &buckets) ||
        __builtin_add_overflow(scalar_bytes, buckets, &total))
        return 0;
    return total;
}

// batch-affine workspace: affine buckets, projective collision buckets, the
// schedule (3 x u32 per entry) and two field arrays of the batch size, each
// array aligned to its element
size_t batch_affine_algorithm_scratch(size_t window, size_t batch, const FrLayout& G) {
    if (window < 1 || window > 24) return 0;
    auto align = [](size_t v, size_t a) { return (v + a - 1) & ~(a - 1); };
    const size_t B = (size_t)1 << (window - 1);
    size_t off = B * G.affine_bucket;
    off = align(off, 8) + B * G.projective_bucket;
    off = align(off, 4) + batch * 12;
    off = align(off, 8) + batch * G.field;
    off = align(off, 8) + batch * G.field;
    return align(off, 8);
}

size_t batch_affine_fr_scratch(size_t n, size_t window, size_t batch, const FrLayout& G = kFrG1) {
    if (n == 0 || n > 0xffffffffULL) return 0;
    if (batch != 128 && batch != 384) return 0;
    const size_t alg = batch_affine_algorithm_scratch(window, batch, G);
    if (alg == 0) return 0;
    size_t scalar_bytes, total;
    if (__builtin_mul_overflow(n, (size_t)32, &scalar_bytes)) return 0;
    if (n == 1) return scalar_bytes;
    if (__builtin_add_overflow(scalar_bytes, alg, &total)) return 0;
    return total;
}

// the reference's algorithm choice for blst_p1s_mult_fr / blst_p2s_mult_fr:
// batch-affine from 2^12 points to the group's measured maximum (2^18 for
// G1, 2^14 for G2; window by size, batch 384 from 2^18), else Pippenger
bool msm_fr_batch_plan(size_t n, size_t& window, size_t& batch, const FrLayout& G = kFrG1) {
    if (n < ((size_t)1 << 12) || n > G.batch_affine_max) return false;
    if (n <= ((size_t)1 << 12)) window = 10;
    else if (n <= ((size_t)1 << 14)) window = 11;
    else if (n <= ((size_t)1 << 15)) window = 12;
    else if (n <= ((size_t)1 << 16)) window = 13;
    else if (n <= ((size_t)1 << 17)) window = 14;
    else window = 15;
    batch = n >= ((size_t)1 << 18) ? 384 : 128;
    return true;
}

size_t msm_fr_scratch(size_t max_points, const FrLayout& G = kFrG1) {
    if (max_points == 0) return 0;
    size_t required = pippenger_fr_scratch(max_points, G);
    if (required == 0) return 0;
    size_t window, batch;
    if (msm_fr_batch_plan(max_points, window, batch, G)) {
        const size_t b = batch_affine_fr_scratch(max_points, window, batch, G);
        if (b == 0) return 0;
        required = std::max(required, b);
    }
    // a capacity query: storage for max_points also runs every smaller
    // input, including the largest batch-affine plan
    const size_t measured_max = G.batch_affine_max;
    if (max_points > measured_max) {
        msm_fr_batch_plan(measured_max, window, batch, G);
        const size_t b = batch_affine_fr_scratch(measured_max, window, batch, G);
        if (b == 0) return 0;
        required = std::max(required, b);
    }
    return required;
}

// the device MSM of contiguous points and Montgomery-form Fr scalars; ret is
// written only on success
bool fr_msm(blst_p1* ret, const blst_p1_affine points[], const void* fr_scalars, size_t n) {
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        blst_p1 r;
        host_buffers_msm(cx, reinterpret_cast<const uint8_t*>(points), static_cast<const uint8_t*>(fr_scalars), n,
                         255, 0, true, true, 0, &r);
        *ret = r;
    } catch (...) {
        g_err = current_error();
        return false;
    }
    g_err.clear();
    return true;
}

bool fr_msm(blst_p2* ret, const blst_p2_affine points[], const void* fr_scalars, size_t n) {
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        blst_p2 r;
        host_buffers_msm_g2(cx, reinterpret_cast<const uint8_t*>(points), static_cast<const uint8_t*>(fr_scalars), n,
                            255, 0, true, true, 0, &r);
        *ret = r;
    } catch (...) {
        g_err = current_error();
        return false;
    }
    g_err.clear();
    return true;
}

// checks of c_api_pippenger_fr / c_api_msm_fr (c_api.cpp:532-552, 809-828),
// the same for both groups
template <class Ret, class Aff>
bool fr_entry(Ret* ret, const Aff points[], const void* scalars, size_t n, blst_limb_t* scratch,
              size_t scratch_size, size_t required) {
    if (ret == nullptr) {
        g_err = "ret is NULL";
        return false;
    }
    if (n == 0) {
        std::memset(ret, 0, sizeof(*ret));
        g_err.clear();
        return true;
    }
    if (required == 0 || points == nullptr || scalars == nullptr || scratch == nullptr || scratch_size < required) {
        g_err = "invalid arguments";
        return false;
    }
    if (n >= 0x7fffffffULL) {   // point indices are 31-bit on the device
        g_err = "npoints beyond 2^31 - 2";
        return false;
    }
    return fr_msm(ret, points, scalars, n);
}

}  // namespace

extern "C" {

size_t blst_p1s_mult_pippenger_scratch_sizeof(size_t npoints) {
    if (npoints == 0) return 0;
    return (size_t)192 << (ref_window_size(npoints) - 1);
}

size_t blst_p1s_mult_pippenger_window_scratch_sizeof(size_t window) {
    if (window < 1 || window > 24) return 0;
    return (size_t)192 << (window - 1);
}

size_t blst_p1s_mult_pippenger_fr_scratch_sizeof(size_t npoints) { return pippenger_fr_scratch(npoints); }

bool blst_p1s_mult_pippenger_fr(blst_p1* ret, const blst_p1_affine points[], const blst_fr scalars[], size_t npoints,
                                blst_limb_t* scratch, size_t scratch_size) {
    return fr_entry(ret, points, scalars, npoints, scratch, scratch_size, pippenger_fr_scratch(npoints));
}

size_t blst_p1s_mult_fr_scratch_sizeof(size_t max_points) { return msm_fr_scratch(max_points); }

bool blst_p1s_mult_fr(blst_p1* ret, const blst_p1_affine points[], const blst_fr scalars[], size_t npoints,
                      blst_limb_t* scratch, size_t scratch_size) {
    return fr_entry(ret, points, scalars, npoints, scratch, scratch_size, msm_fr_scratch(npoints));
}

void blst_p1s_mult_pippenger(blst_p1* ret, const blst_p1_affine* const points[], size_t npoints,
                             const blst_byte* const scalars[], size_t nbits, blst_limb_t* scratch) {
    if (npoints == 0 || nbits == 0) { std::memset(ret, 0, sizeof(*ret)); g_err.clear(); return; }
    if (nbits > 256 || !heads_ok((const void* const*)points, (const void* const*)scalars, npoints)) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "invalid arguments";
        return;
    }
    if (npoints > 1 && scratch == nullptr) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "scratch is NULL";
        return;
    }
    host_msm(ret, points, npoints, scalars, nbits, 0, true, 0);
}

void blst_p1s_mult_pippenger_window(blst_p1* ret, const blst_p1_affine* const points[], size_t npoints,
                                    const blst_byte* const scalars[], size_t nbits,
                                    blst_limb_t* scratch, size_t window) {
    if (npoints == 0 || nbits == 0) { std::memset(ret, 0, sizeof(*ret)); g_err.clear(); return; }
    if (nbits > 256 || window < 1 || window > 24 ||
        !heads_ok((const void* const*)points, (const void* const*)scalars, npoints)) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "invalid arguments";
        return;
    }
    if (npoints > 1 && scratch == nullptr) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "scratch is NULL";
        return;
    }
    host_msm(ret, points, npoints, scalars, nbits, window, true, 0);
}

void blst_p1s_tile_pippenger(blst_p1* ret, const blst_p1_affine* const points[], size_t npoints,
                             const blst_byte* const scalars[], size_t nbits, blst_limb_t* scratch,
                             size_t bit0, size_t window) {
    if (npoints == 0 || !heads_ok((const void* const*)points, (const void* const*)scalars, npoints) ||
        scratch == nullptr || window < 1 || window > 24 || nbits == 0 || nbits > 256 || bit0 >= nbits) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "invalid arguments";
        return;
    }
    host_msm(ret, points, npoints, scalars, nbits, window, false, bit0);
}

size_t blst_p2s_mult_pippenger_scratch_sizeof(size_t npoints) {
    if (npoints == 0) return 0;
    return (size_t)384 << (ref_window_size(npoints) - 1);
}

size_t blst_p2s_mult_pippenger_window_scratch_sizeof(size_t window) {
    if (window < 1 || window > 24) return 0;
    return (size_t)384 << (window - 1);
}

size_t blst_p2s_mult_pippenger_fr_scratch_sizeof(size_t npoints) { return pippenger_fr_scratch(npoints, kFrG2); }

bool blst_p2s_mult_pippenger_fr(blst_p2* ret, const blst_p2_affine points[], const blst_fr scalars[], size_t npoints,
                                blst_limb_t* scratch, size_t scratch_size) {
    return fr_entry(ret, points, scalars, npoints, scratch, scratch_size, pippenger_fr_scratch(npoints, kFrG2));
}

size_t blst_p2s_mult_fr_scratch_sizeof(size_t max_points) { return msm_fr_scratch(max_points, kFrG2); }

bool blst_p2s_mult_fr(blst_p2* ret, const blst_p2_affine points[], const blst_fr scalars[], size_t npoints,
                      blst_limb_t* scratch, size_t scratch_size) {
    return fr_entry(ret, points, scalars, npoints, scratch, scratch_size, msm_fr_scratch(npoints, kFrG2));
}

void blst_p2s_mult_pippenger(blst_p2* ret, const blst_p2_affine* const points[], size_t npoints,
                             const blst_byte* const scalars[], size_t nbits, blst_limb_t* scratch) {
    if (npoints == 0 || nbits == 0) { std::memset(ret, 0, sizeof(*ret)); g_err.clear(); return; }
    if (nbits > 256 || !heads_ok((const void* const*)points, (const void* const*)scalars, npoints) ||
        npoints >= 0x7fffffffULL) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "invalid arguments";
        return;
    }
    if (npoints > 1 && scratch == nullptr) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "scratch is NULL";
        return;
    }
    host_msm_g2(ret, points, npoints, scalars, nbits, 0, true, 0);
}

void blst_p2s_mult_pippenger_window(blst_p2* ret, const blst_p2_affine* const points[], size_t npoints,
                                    const blst_byte* const scalars[], size_t nbits,
                                    blst_limb_t* scratch, size_t window) {
    if (npoints == 0 || nbits == 0) { std::memset(ret, 0, sizeof(*ret)); g_err.clear(); return; }
    if (nbits > 256 || window < 1 || window > 24 || npoints >= 0x7fffffffULL ||
        !heads_ok((const void* const*)points, (const void* const*)scalars, npoints)) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "invalid arguments";
        return;
    }
    if (npoints > 1 && scratch == nullptr) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "scratch is NULL";
        return;
    }
    host_msm_g2(ret, points, npoints, scalars, nbits, window, true, 0);
}

void blst_p2s_tile_pippenger(blst_p2* ret, const blst_p2_affine* const points[], size_t npoints,
                             const blst_byte* const scalars[], size_t nbits, blst_limb_t* scratch,
                             size_t bit0, size_t window) {
    if (npoints == 0 || !heads_ok((const void* const*)points, (const void* const*)scalars, npoints) ||
        scratch == nullptr || window < 1 || window > 24 || nbits == 0 || nbits > 256 || bit0 >= nbits ||
        npoints >= 0x7fffffffULL) {
        std::memset(ret, 0, sizeof(*ret));
        g_err = "invalid arguments";
        return;
    }
    host_msm_g2(ret, points, npoints, scalars, nbits, window, false, bit0);
}

int blst_hip_p2_msm_device(blst_p2* ret, const void* device_points, const void* device_scalars, size_t npoints,
                           size_t nbits, size_t window) {
    if (!ret) { g_err = "invalid arguments"; return -1; }
    std::memset(ret, 0, sizeof(*ret));
    if (npoints == 0 || nbits == 0) { g_err.clear(); return 0; }
    if (!device_points || !device_scalars || nbits > 256 || window > 24 || npoints >= 0x7fffffffULL ||
        ((uintptr_t)device_points % 16)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        JobSpec js = make_spec(device_points, device_scalars, npoints, nbits, window, false);
        device_msm_g2(cx, js, ret);
    } catch (...) {
        g_err = current_error();
        std::memset(ret, 0, sizeof(*ret));
        return -1;
    }
    g_err.clear();
    return 0;
}

const char* blst_hip_last_error(void) { return g_err.c_str(); }

int blst_hip_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return 0;
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int blst_hip_msm_device(blst_p1* ret, const void* device_points, const void* device_scalars,
                        size_t npoints, size_t nbits, size_t window) {
    return blst_hip_msm_device_tuned(ret, device_points, device_scalars, npoints, nbits, window, 0, 0);
}

int blst_hip_msm_device_tuned(blst_p1* ret, const void* device_points, const void* device_scalars,
                              size_t npoints, size_t nbits, size_t window, size_t chunk, size_t segment) {
    auto t0 = std::chrono::steady_clock::now();
    std::memset(ret, 0, sizeof(*ret));
    if (npoints == 0 || nbits == 0) { g_err.clear(); return 0; }
    if (nbits > 256 || window > 24 || chunk > 4096 || segment > 4096 || !device_points || !device_scalars ||
        ((uintptr_t)device_points % 16) != 0 || npoints >= 0x7fffffffULL) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        JobSpec js = make_spec(device_points, device_scalars, npoints, nbits, window, false);
        if (chunk) js.T = (uint32_t)chunk;
        if (segment) js.L = (uint32_t)segment;
        // large MSMs run in point ranges (each range's sort scatters its
        // entries within a few tens of MB); a caller's tile keeps one range
        const uint32_t K = chunk ? 1 : point_ranges(npoints, kDeviceSplitMin, kDeviceRangePoints);
        if (K > 1) ranged_msm(cx, js, (nbits + 7) / 8, K, [](uint32_t, size_t, size_t) {}, ret);
        else device_msm(cx, js, ret);
    } catch (...) {
        g_err = current_error();
        std::memset(ret, 0, sizeof(*ret));
        return -1;
    }
    g_phase[6] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    g_err.clear();
    return 0;
}

int blst_hip_msm_device_noalloc(blst_p1* ret, const void* device_points, const void* device_scalars,
                                size_t npoints, size_t nbits) {
    if (!ret) {
        g_err = "invalid arguments";
        return -1;
    }
    std::memset(ret, 0, sizeof(*ret));
    if (npoints == 0 || nbits == 0) { g_err.clear(); return 0; }
    if (nbits > 256 || !device_points || !device_scalars || ((uintptr_t)device_points % 16) != 0) {
        g_err = "invalid arguments";
        return -1;
    }
    if (npoints > kNoAllocMaxPoints) {
        g_err = "npoints above the " + std::to_string(kNoAllocMaxPoints) + "-point limit of the no-allocation MSM";
        return -1;
    }
    try {
        noalloc_msm(device_points, device_scalars, npoints, nbits, ret);
    } catch (...) {
        g_err = current_error();
        std::memset(ret, 0, sizeof(*ret));
        return -1;
    }
    g_err.clear();
    return 0;
}

int blst_hip_msm_batch_device(blst_p1* rets, const blst_hip_msm_job* jobs, size_t njobs) {
    auto t0 = std::chrono::steady_clock::now();
    if (njobs == 0) { g_err.clear(); return 0; }
    if (!rets || !jobs) { g_err = "invalid arguments"; return -1; }
    std::memset(rets, 0, njobs * sizeof(blst_p1));
    std::vector<JobSpec> specs;
    std::vector<size_t> idx;
    if (!specs_from_jobs(jobs, njobs, specs, idx)) return -1;
    if (specs.empty()) { g_err.clear(); return 0; }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        Layout L = make_layout(specs);
        std::vector<hostec::xyzz6> res;
        run_batch(cx.ws, L, false, res);
        auto t1 = std::chrono::steady_clock::now();
        for (size_t k = 0; k < specs.size(); k++) {
            hostec::jac6 jj;
            hostec::to_jac(jj, res[k]);
            std::memcpy(&rets[idx[k]], &jj, sizeof(jj));
        }
        g_phase[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    } catch (...) {
        g_err = current_error();
        std::memset(rets, 0, njobs * sizeof(blst_p1));
        return -1;
    }
    g_phase[6] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    g_err.clear();
    return 0;
}

blst_hip_batch_plan* blst_hip_batch_plan_create(const blst_hip_msm_job* jobs, size_t njobs) {
    if (!jobs && njobs) { g_err = "invalid arguments"; return nullptr; }
    auto plan = std::make_unique<blst_hip_batch_plan>();
    PlanImpl& P = plan->impl;
    P.njobs = njobs;
    std::vector<JobSpec> specs;
    if (!specs_from_jobs(jobs, njobs, specs, P.idx)) return nullptr;
    try {
        HIPCK(hipGetDevice(&P.device));
        P.ws.init();
        if (!specs.empty()) {
            P.L = make_layout(specs);
            ensure_workspace(P.ws, P.L);
            P.D = upload_layout(P.ws, P.L);
            HIPCK(hipHostMalloc((void**)&P.host_res, specs.size() * sizeof(hostec::xyzz6), hipHostMallocDefault));
            HIPCK(hipStreamSynchronize(P.ws.stream));
            HIPCK(hipStreamBeginCapture(P.ws.stream, hipStreamCaptureModeThreadLocal));
            enqueue_batch(P.ws, P.L, P.D, false, P.host_res, false);
            HIPCK(hipStreamEndCapture(P.ws.stream, &P.graph));
            HIPCK(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
        }
    } catch (...) {
        g_err = current_error();
        return nullptr;
    }
    g_err.clear();
    return plan.release();
}

int blst_hip_batch_plan_run(blst_hip_batch_plan* plan, blst_p1* rets) {
    auto t0 = std::chrono::steady_clock::now();
    if (!plan || !rets) { g_err = "invalid arguments"; return -1; }
    PlanImpl& P = plan->impl;
    std::memset(rets, 0, P.njobs * sizeof(blst_p1));
    if (P.idx.empty()) { g_err.clear(); return 0; }
    try {
        HIPCK(hipGraphLaunch(P.exec, P.ws.stream));
        HIPCK(hipStreamSynchronize(P.ws.stream));
        for (size_t k = 0; k < P.idx.size(); k++) {
            hostec::jac6 jj;
            hostec::to_jac(jj, P.host_res[k]);
            std::memcpy(&rets[P.idx[k]], &jj, sizeof(jj));
        }
    } catch (...) {
        g_err = current_error();
        return -1;
    }
    g_phase[6] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    g_err.clear();
    return 0;
}

void blst_hip_batch_plan_destroy(blst_hip_batch_plan* plan) { delete plan; }

void* blst_hip_device_upload(const void* host, size_t bytes) {
    void* d = nullptr;
    if (!host || bytes == 0) { g_err = "invalid arguments"; return nullptr; }
    if (hipMalloc(&d, bytes) != hipSuccess) { g_err = "hipMalloc failed"; return nullptr; }
    if (hipMemcpy(d, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        g_err = "hipMemcpy failed";
        return nullptr;
    }
    g_err.clear();
    return d;
}

int blst_hip_device_write(void* device_dst, const void* host_src, size_t bytes) {
    if (!device_dst || !host_src) { g_err = "invalid arguments"; return -1; }
    if (hipMemcpy(device_dst, host_src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        g_err = "hipMemcpy failed";
        return -1;
    }
    g_err.clear();
    return 0;
}

void* blst_hip_host_alloc(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0 || hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        g_err = "hipHostMalloc failed";
        return nullptr;
    }
    g_err.clear();
    return p;
}

void blst_hip_host_free(void* host_ptr) {
    if (host_ptr) (void)hipHostFree(host_ptr);
}

void* blst_hip_device_alloc(size_t bytes) {
    void* d = nullptr;
    if (bytes == 0) { g_err = "invalid arguments"; return nullptr; }
    if (hipMalloc(&d, bytes) != hipSuccess) { g_err = "hipMalloc failed"; return nullptr; }
    g_err.clear();
    return d;
}

int blst_hip_device_read(void* host_dst, const void* device_src, size_t bytes) {
    if (!host_dst || !device_src) { g_err = "invalid arguments"; return -1; }
    if (hipMemcpy(host_dst, device_src, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
        g_err = "hipMemcpy failed";
        return -1;
    }
    g_err.clear();
    return 0;
}

void blst_hip_device_free(void* device_ptr) {
    if (device_ptr) (void)hipFree(device_ptr);
}

// Host buffers both ways: the Jacobian points go up (144 B a point) and the
// affine points come back (96 B), each direction about as long as the other
// and both far longer than the conversion.  From kAffinePipeMin points the
// points move in ranges of kAffineRange: the calling thread stages range r up
// (copies on cx.stream, the conversion on a pipeline stream behind them, so
// the next range's copies do not wait for it), while a second thread brings
// converted ranges down through its own page-locked ring and copy pool on a
// stream of its own: the two directions of the link run at once.
constexpr size_t kAffinePipeMin = (size_t)1 << 19;
constexpr size_t kAffineRange = (size_t)1 << 18;

void host_to_affine(Ctx& cx, blst_p1_affine* dst, const uint8_t* hp, size_t n) {
    const size_t K = (n + kAffineRange - 1) / kAffineRange;
    if (!cx.d2h) HIPCK(hipStreamCreateWithFlags(&cx.d2h, hipStreamNonBlocking));
    if (cx.pipe.empty()) {
        cx.pipe.emplace_back(new Workspace());
        cx.pipe.back()->init();
    }
    hipStream_t conv = cx.pipe[0]->stream;
    // range r: event 2r when its points are up, 2r + 1 when it is converted
    while (cx.range_ready.size() < 2 * K) {
        hipEvent_t e;
        HIPCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        cx.range_ready.push_back(e);
    }
    uint8_t* in = static_cast<uint8_t*>(cx.jac_in.p);
    uint8_t* out = static_cast<uint8_t*>(cx.aff_out.p);
    std::mutex mu;
    std::condition_variable cv;
    size_t queued = 0;          // ranges whose conversion is queued (event recorded)
    bool failed = false;        // the calling thread stopped early
    std::string down_err;
    std::thread downloader([&] {
        try {
            for (size_t r = 0; r < K; r++) {
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return queued > r || failed; });
                    if (queued <= r) return;
                }
                const size_t lo = r * kAffineRange, m = std::min(kAffineRange, n - lo);
                HIPCK(hipStreamWaitEvent(cx.d2h, cx.range_ready[2 * r + 1], 0));
                cx.down.download(dst + lo, out + lo * sizeof(blst_p1_affine), m * sizeof(blst_p1_affine), cx.d2h);
            }
        } catch (...) {
            down_err = current_error();
        }
    });
    try {
        for (size_t r = 0; r < K; r++) {
            const size_t lo = r * kAffineRange, m = std::min(kAffineRange, n - lo);
            cx.stager.upload(in + lo * sizeof(blst_p1), hp + lo * sizeof(blst_p1), m * sizeof(blst_p1), cx.stream);
            HIPCK(hipEventRecord(cx.range_ready[2 * r], cx.stream));
            HIPCK(hipStreamWaitEvent(conv, cx.range_ready[2 * r], 0));
            enqueue_to_affine(conv, in + lo * sizeof(blst_p1), out + lo * sizeof(blst_p1_affine), m);
            HIPCK(hipEventRecord(cx.range_ready[2 * r + 1], conv));
            {
                std::lock_guard<std::mutex> lk(mu);
                queued = r + 1;
            }
            cv.notify_one();
        }
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(mu);
            failed = true;
        }
        cv.notify_one();
        downloader.join();
        throw;
    }
    downloader.join();
    if (!down_err.empty()) throw HipFail{down_err};
}

void blst_p1s_to_affine(blst_p1_affine dst[], const blst_p1* const points[], size_t npoints) {
    if (npoints == 0) { g_err.clear(); return; }
    if (dst == nullptr) { g_err = "dst is NULL"; return; }
    if (points == nullptr || points[0] == nullptr) {
        std::memset(dst, 0, npoints * sizeof(blst_p1_affine));
        g_err.clear();
        return;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        std::vector<uint8_t> stage;
        const uint8_t* hp = resolve(points, npoints, sizeof(blst_p1), stage);
        const size_t chunk = std::min(npoints, kAffineChunk);
        cx.jac_in.ensure(chunk * sizeof(blst_p1));
        cx.aff_out.ensure(chunk * sizeof(blst_p1_affine));
        for (size_t b = 0; b < npoints; b += chunk) {
            const size_t m = std::min(chunk, npoints - b);
            if (m >= kAffinePipeMin) {
                host_to_affine(cx, dst + b, hp + b * sizeof(blst_p1), m);
                continue;
            }
            cx.stager.upload(cx.jac_in.p, hp + b * sizeof(blst_p1), m * sizeof(blst_p1), cx.stream);
            enqueue_to_affine(cx.stream, cx.jac_in.p, cx.aff_out.p, m);
            cx.stager.download(dst + b, cx.aff_out.p, m * sizeof(blst_p1_affine), cx.stream);
        }
    } catch (...) {
        g_err = current_error();
        std::memset(dst, 0, npoints * sizeof(blst_p1_affine));
        return;
    }
    g_err.clear();
}

int blst_hip_p1s_to_affine_device(void* device_dst, const void* device_points, size_t npoints) {
    if (npoints == 0) { g_err.clear(); return 0; }
    if (!device_dst || !device_points || npoints >= 0x7fffffffULL || ((uintptr_t)device_dst % 16) ||
        ((uintptr_t)device_points % 16)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        enqueue_to_affine(cx.stream, device_points, device_dst, npoints);
        HIPCK(hipStreamSynchronize(cx.stream));
    } catch (...) {
        g_err = current_error();
        return -1;
    }
    g_err.clear();
    return 0;
}

size_t blst_p1s_mult_wbits_precompute_sizeof(size_t wbits, size_t npoints) {
    return wbits_precompute_bytes<G1Ops>(wbits, npoints);
}

size_t blst_p1s_mult_wbits_scratch_sizeof(size_t npoints) {
    return 144 * std::min<size_t>(npoints, 8192);
}

void blst_p1s_mult_wbits_precompute(blst_p1_affine table[], size_t wbits, const blst_p1_affine* const points[],
                                    size_t npoints) {
    wbits_precompute_entry<G1Ops>(table, wbits, points, npoints);
}

void blst_p1s_mult_wbits(blst_p1* ret, const blst_p1_affine table[], size_t wbits, size_t npoints,
                         const blst_byte* const scalars[], size_t nbits, blst_limb_t* scratch) {
    mult_wbits_entry<G1Ops>(ret, table, wbits, npoints, scalars, nbits, scratch);
}

size_t blst_p2s_mult_wbits_precompute_sizeof(size_t wbits, size_t npoints) {
    return wbits_precompute_bytes<G2Ops>(wbits, npoints);
}

// one BatchAffineNode<Fp2> (affine point + Fp2, 288 bytes) per point, at
// most fixed_table_scratch_capacity<Fp2>() = 4096 (multi_scalar.h:218-222,
// 310-315)
size_t blst_p2s_mult_wbits_scratch_sizeof(size_t npoints) {
    return 288 * std::min<size_t>(npoints, 4096);
}

void blst_p2s_mult_wbits_precompute(blst_p2_affine table[], size_t wbits, const blst_p2_affine* const points[],
                                    size_t npoints) {
    wbits_precompute_entry<G2Ops>(table, wbits, points, npoints);
}

void blst_p2s_mult_wbits(blst_p2* ret, const blst_p2_affine table[], size_t wbits, size_t npoints,
                         const blst_byte* const scalars[], size_t nbits, blst_limb_t* scratch) {
    mult_wbits_entry<G2Ops>(ret, table, wbits, npoints, scalars, nbits, scratch);
}

void blst_p2s_to_affine(blst_p2_affine dst[], const blst_p2* const points[], size_t npoints) {
    p2s_to_affine(dst, points, npoints);
}

int blst_hip_p1s_mult_wbits_precompute_device(void* device_table, size_t wbits, const void* device_points,
                                              size_t npoints) {
    if (npoints == 0) { g_err.clear(); return 0; }
    if (!device_table || !device_points || !valid_wbits(wbits) || npoints >= 0x7fffffffULL ||
        ((uintptr_t)device_table % 16) || ((uintptr_t)device_points % 16)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        const size_t m = (size_t)1 << (wbits - 1);
        const size_t R = std::max<size_t>(1, kWbitsRowEntries / m);
        for (size_t row0 = 0; row0 < npoints; row0 += R) {
            const size_t nr = std::min(R, npoints - row0);
            enqueue_wbits_table(cx, static_cast<const uint8_t*>(device_points) + row0 * sizeof(blst_p1_affine), nr,
                                wbits, static_cast<uint8_t*>(device_table) + row0 * m * sizeof(blst_p1_affine));
        }
        HIPCK(hipStreamSynchronize(cx.stream));
    } catch (...) {
        g_err = current_error();
        return -1;
    }
    g_err.clear();
    return 0;
}

int blst_hip_p1s_mult_wbits_device(blst_p1* ret, const void* device_table, size_t wbits, size_t npoints,
                                   const void* device_scalars, size_t nbits) {
    if (!ret) { g_err = "invalid arguments"; return -1; }
    std::memset(ret, 0, sizeof(*ret));
    if (npoints == 0 || nbits == 0) { g_err.clear(); return 0; }
    if (!device_table || !device_scalars || !valid_wbits(wbits) || nbits > 256 || npoints >= 0x7fffffffULL ||
        ((uintptr_t)device_table % 16)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        std::vector<hostec::xyzz6> W;
        wbits_window_sums(cx, device_table, device_scalars, nbits, nullptr, npoints, wbits, W);
        hostec::xyzz6 r;
        hostec::fold_windows(r, W.data(), (int)W.size(), (int)wbits);
        hostec::jac6 j;
        hostec::to_jac(j, r);
        std::memcpy(ret, &j, sizeof(j));
    } catch (...) {
        g_err = current_error();
        std::memset(ret, 0, sizeof(*ret));
        return -1;
    }
    g_err.clear();
    return 0;
}

size_t blst_hip_comb_table_bytes(size_t npoints, size_t wbits, size_t nbits) {
    return comb_table_bytes(npoints, wbits, nbits);
}

int blst_hip_comb_table_build(void* device_table, size_t wbits, size_t nbits, const void* device_points,
                              size_t npoints) {
    if (npoints == 0) { g_err.clear(); return 0; }
    if (!device_table || !device_points || comb_table_bytes(npoints, wbits, nbits) == 0 ||
        npoints >= 0x7fffffffULL || ((uintptr_t)device_table % 16) || ((uintptr_t)device_points % 16)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        const size_t D = 1 + nbits / wbits, m = (size_t)1 << (wbits - 1);
        cx.comb_q.ensure(npoints * sizeof(blst_p1_affine));
        cx.comb_x.ensure(npoints * sizeof(xyzz));
        HIPCK(hipMemcpyAsync(cx.comb_q.p, device_points, npoints * sizeof(blst_p1_affine), hipMemcpyDeviceToDevice,
                             cx.stream));
        const size_t R = std::max<size_t>(1, kWbitsRowEntries / m);
        for (size_t w = 0; w < D; w++) {
            uint8_t* slice = static_cast<uint8_t*>(device_table) + w * npoints * m * sizeof(blst_p1_affine);
            for (size_t row0 = 0; row0 < npoints; row0 += R) {
                const size_t nr = std::min(R, npoints - row0);
                enqueue_wbits_table(cx, cx.comb_q.as<uint8_t>() + row0 * sizeof(blst_p1_affine), nr, wbits,
                                    slice + row0 * m * sizeof(blst_p1_affine));
            }
            if (w + 1 < D) {
                // the next window's bases: [2^wbits] Q
                launch_k_shift_points(cx.stream, cx.comb_q.as<const uint4>(), (uint32_t)npoints, (uint32_t)wbits,
                                      cx.comb_x.as<xyzz>());
                const uint32_t Q = (uint32_t)std::min<size_t>(64, std::max<size_t>(1, npoints / kTableAffineLanes));
                launch_k_to_affine(dim3(div_up(div_up(npoints, Q), 256)), dim3(256), cx.stream,
                                   cx.comb_x.as<const xyzz>(), cx.comb_q.as<aff>(), (uint32_t)npoints, Q);
                HIPCK(hipGetLastError());
            }
        }
        HIPCK(hipStreamSynchronize(cx.stream));
    } catch (...) {
        g_err = current_error();
        return -1;
    }
    g_err.clear();
    return 0;
}

int blst_hip_comb_msm_batch(blst_p1* rets, const void* device_table, size_t wbits, size_t npoints,
                            const void* device_scalars, size_t njobs, size_t nbits) {
    if (njobs == 0) { g_err.clear(); return 0; }
    if (!rets) { g_err = "invalid arguments"; return -1; }
    std::memset(rets, 0, njobs * sizeof(blst_p1));
    if (npoints == 0 || nbits == 0) { g_err.clear(); return 0; }
    if (!device_table || !device_scalars || comb_table_bytes(npoints, wbits, nbits) == 0 ||
        npoints >= 0x7fffffffULL || njobs >= 0x7fffffffULL || ((uintptr_t)device_table % 16)) {
        g_err = "invalid arguments";
        return -1;
    }
    try {
        Ctx& cx = context();
        std::lock_guard<std::mutex> lk(cx.mu);
        const size_t D = 1 + nbits / wbits, nbytes = (nbits + 7) / 8;
        const size_t terms = D * npoints;
        if (terms >= 0xffffffffULL) throw std::runtime_error("too many table terms");
        const size_t C = std::min(terms, std::max<size_t>(4, (terms * njobs + 4 * kWbitsSumLanes - 1) /
                                                                 (4 * kWbitsSumLanes)));
        const size_t nch = (terms + C - 1) / C;
        if (nch * njobs >= 0xffffffffULL) throw std::runtime_error("batch too large");
        cx.wb_part.ensure(njobs * nch * sizeof(xyzz));
        cx.wb_W.ensure(njobs * sizeof(xyzz));
        const uint32_t vec = (nbytes == 32 && ((uintptr_t)device_scalars % 16) == 0) ? 1u : 0u;
        launch_k_comb_sum(cx.stream, static_cast<const uint4*>(device_table),
                          static_cast<const uint8_t*>(device_scalars), (uint32_t)nbytes, (uint32_t)nbits, vec,
                          (uint32_t)npoints, (uint32_t)wbits, (uint32_t)D, (uint32_t)njobs, (uint32_t)C,
                          (uint32_t)nch, cx.wb_part.as<xyzz>());
        launch_k_wbits_tree(cx.stream, cx.wb_part.as<const xyzz>(), (uint32_t)nch, (uint32_t)njobs,
                            cx.wb_W.as<xyzz>());
        HIPCK(hipGetLastError());
        std::vector<hostec::xyzz6> res(njobs);
        HIPCK(hipMemcpyAsync(res.data(), cx.wb_W.p, njobs * sizeof(xyzz), hipMemcpyDeviceToHost, cx.stream));
        HIPCK(hipStreamSynchronize(cx.stream));
        for (size_t k = 0; k < njobs; k++) {
            hostec::jac6 j;
            hostec::to_jac(j, res[k]);
            std::memcpy(&rets[k], &j, sizeof(j));
        }
    } catch (...) {
        g_err = current_error();
        return -1;
    }
    g_err.clear();
    return 0;
}

void blst_hip_p1_compress(blst_byte out[48], const blst_p1* in) {
    hostec::jac6 j;
    std::memcpy(&j, in, sizeof(j));
    hostec::compress(out, j);
}

int blst_hip_tuning_set(int batch, size_t log2n, size_t window, size_t chunk, size_t segment) {
    if (log2n > 31 || window < 1 || window > 24 || chunk < 1 || chunk > 4096 || segment < 1 || segment > 4096) {
        g_err = "invalid arguments";
        return -1;
    }
    set_tuning(batch != 0, (int)log2n, Tune{(uint32_t)window, (uint32_t)chunk, (uint32_t)segment});
    g_err.clear();
    return 0;
}

void blst_hip_tuning_get(int batch, size_t npoints, size_t nbits, size_t out[3]) {
    const Tune t = tuned(npoints, nbits, batch != 0);
    out[0] = t.c;
    out[1] = t.T;
    out[2] = t.L;
}

int blst_hip_set_pipeline(size_t groups) {
    if (groups > 16) {
        g_err = "invalid arguments";
        return -1;
    }
    g_pipeline.store((uint32_t)groups);
    g_err.clear();
    return 0;
}

int blst_hip_set_pipeline_share(size_t percent) {
    if (percent > 50) {
        g_err = "invalid arguments";
        return -1;
    }
    g_last_share.store((uint32_t)percent);
    g_err.clear();
    return 0;
}

void blst_hip_last_phase_ms(double out[7]) {
    for (int k = 0; k < 7; k++) out[k] = g_phase[k];
}

}  // extern "C"
