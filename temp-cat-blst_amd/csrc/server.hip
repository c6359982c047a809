// Serving loop: continuous batching of MSM requests on one GPU.
//
// Clients submit requests (over resident bases such as a KZG setup, or over
// fresh bases that travel with the request) and wait for tickets.  A
// scheduler thread
//   * admits pending requests in arrival order while their device footprint
//     (inputs plus the batch workspace the request adds) fits the HBM budget
//     and the points in flight stay within a few batches,
//   * has a copier thread move the inputs of admitted requests on a copy
//     stream (device buffers from a slab pool),
//   * whenever one of three batch slots is free, launches every admitted
//     request whose copies have landed as one batch (one digit pass, one
//     sort, one accumulation, one reduction over all of them); the slots
//     queue their batches on one compute stream, so the device runs them
//     back to back while copies, batch set-up and completions overlap,
//   * completes finished batches: the window sums come back to the host and
//     a thread pool folds them per request, then the tickets are released.
// Requests that arrive while batches run join the next batch as soon as
// memory and a slot allow.
#include "pipeline.h"
#include "../../include/blst_hip.h"

#include <atomic>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace bls {
namespace host {
namespace {

using clock_t_ = std::chrono::steady_clock;

// Device buffers recycled by capacity: hipFree synchronises the device, so
// the loop never frees while serving.  Capacities are powers of two; classes
// below kSlab are carved from kSlab-sized allocations, so a stream of small
// requests costs one hipMalloc per slab.  Thread-safe.
class BufferPool {
public:
    static constexpr size_t kSlab = (size_t)64 << 20;
    ~BufferPool() {
        for (void* p : blocks_) (void)hipFree(p);
    }
    void* get(size_t bytes, size_t& cap) {
        cap = round_up(bytes);
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.find(cap);
        if (it == free_.end()) {
            const size_t span = std::max(cap, kSlab);
            void* p = nullptr;
            device_malloc(&p, span);
            blocks_.push_back(p);
            allocated_ += span;
            for (size_t o = 0; o + cap <= span; o += cap) free_.emplace(cap, static_cast<uint8_t*>(p) + o);
            it = free_.find(cap);
        }
        void* p = it->second;
        free_.erase(it);
        return p;
    }
    void put(void* p, size_t cap) {
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(cap, p);
    }

private:
    static size_t round_up(size_t b) {
        size_t c = 1 << 16;
        while (c < b) c <<= 1;
        return c;
    }
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    std::vector<void*> blocks_;
    size_t allocated_ = 0;
};

// Host threads for the window folds of finished batches (sum_w 2^(cw) W_w
// per request: a chain of ~255 doublings, cheap on a CPU core and latency
// bound on the GPU).  run() hands out indices to the pool and the caller.
class FoldPool {
public:
    FoldPool() {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned n = std::max(1u, std::min(8u, hw / 2));
        for (unsigned i = 0; i + 1 < n; i++) workers_.emplace_back([this] { loop(); });
    }
    ~FoldPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    void run(size_t n, std::function<void(size_t)> f) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            f_ = std::move(f);
            n_ = n;
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
        for (size_t i; (i = next_.fetch_add(1)) < n_;) f_(i);
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
    std::function<void(size_t)> f_;
    size_t n_ = 0, pending_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
};

struct Request {
    int64_t ticket = 0;
    int base_id = -1;
    const void* host_points = nullptr;
    const void* host_scalars = nullptr;
    size_t n = 0, nbits = 0;
    size_t need = 0;                  // admitted footprint
    void* d_points = nullptr;         // fresh bases only
    size_t d_points_cap = 0;
    void* d_scalars = nullptr;
    size_t d_scalars_cap = 0;
    hipEvent_t copied = nullptr;
    bool copy_issued = false;         // set by the copier thread
    bool copying = false;             // the copier works on it with mu_ released
    bool failed = false;              // failed while copying: the copier releases it
    bool landed = false;              // copies complete (seen by the scheduler)
    clock_t_::time_point t_submit;
    bool done = false;
    int status = 0;
    blst_p1 result = {};
    double latency_ms = 0;
};

// batches in flight at once: one finishing (results, merges, folds), one
// accumulating, one whose inputs have landed and that is ready to launch
constexpr int kSlots = 3;
constexpr int kQueued = 2;

struct Slot {
    Workspace ws;
    bool busy = false;
    Layout L;
    std::vector<std::shared_ptr<Request>> reqs;
    hostec::xyzz6* host_res = nullptr;
    size_t host_res_cap = 0;
    hipEvent_t done = nullptr;
    bool gpu_done = false;            // batch finished on the device, not yet completed
};

}  // namespace

class Server {
public:
    Server(size_t budget, size_t max_batch_points) : max_batch_points_(max_batch_points) {
        HIPCK(hipGetDevice(&device_));
        size_t free_b = 0, total_b = 0;
        HIPCK(hipMemGetInfo(&free_b, &total_b));
        budget_ = budget ? budget : (size_t)(0.8 * (double)free_b);
        HIPCK(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
        // one compute stream for all slots: batches run one after another on
        // the device (concurrent batches measured slower than batches in
        // turn), while copies, batch set-up and completions overlap them
        HIPCK(hipStreamCreateWithFlags(&compute_stream_, hipStreamNonBlocking));
        for (auto& s : slots_) {
            s = std::make_unique<Slot>();
            s->ws.init(compute_stream_);
            s->ws.kernel_transfers = true;
            HIPCK(hipEventCreateWithFlags(&s->done, hipEventDisableTiming));
        }
        worker_ = std::thread([this] { loop(); });
        copier_ = std::thread([this] { copy_loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        copy_cv_.notify_all();
        worker_.join();
        copier_.join();
        for (auto& s : slots_) {
            if (s->host_res) (void)hipHostFree(s->host_res);
            if (s->done) (void)hipEventDestroy(s->done);
        }
        for (auto& b : bases_) (void)hipFree(b.ptr);
        (void)hipStreamDestroy(copy_stream_);
        if (compute_stream_) (void)hipStreamDestroy(compute_stream_);
    }

    int add_bases(const blst_p1_affine* pts, size_t n) {
        std::lock_guard<std::mutex> lk(mu_);
        Base b;
        b.n = n;
        HIPCK(hipSetDevice(device_));
        device_malloc(&b.ptr, n * sizeof(blst_p1_affine));
        HIPCK(hipMemcpy(b.ptr, pts, n * sizeof(blst_p1_affine), hipMemcpyHostToDevice));
        used_ += n * sizeof(blst_p1_affine);
        bases_.push_back(b);
        return (int)bases_.size() - 1;
    }

    int64_t submit(int base_id, const blst_p1_affine* pts, const blst_byte* scalars, size_t n, size_t nbits,
                   std::string& err) {
        auto r = std::make_shared<Request>();
        r->base_id = base_id;
        r->host_points = pts;
        r->host_scalars = scalars;
        r->n = n;
        r->nbits = nbits;
        r->t_submit = clock_t_::now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (base_id >= 0 && ((size_t)base_id >= bases_.size() || n > bases_[base_id].n)) {
                err = "unknown bases, or more points than the bases hold";
                return -1;
            }
            // a batch holds every job's (window, point) entries, keys and
            // points as 31-bit indices: a request beyond that could never run
            const uint32_t c = make_spec(nullptr, nullptr, n, nbits, 0, true).c;
            if ((uint64_t)(1 + nbits / c) * n >= 0x7fffffffULL) {
                err = "request too large for one launch (31-bit entry indices)";
                return -1;
            }
            r->need = footprint(*r);
            if (r->need > budget_) {
                err = "request larger than the HBM budget";
                return -1;
            }
            r->ticket = ++next_ticket_;
            tickets_[r->ticket] = r;
            pending_.push_back(r);
            submitted_++;
        }
        cv_.notify_all();
        return r->ticket;
    }

    int wait(int64_t ticket, blst_p1* ret, double* latency_ms, std::string& err) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = tickets_.find(ticket);
        if (it == tickets_.end()) {
            err = "unknown ticket";
            return -1;
        }
        std::shared_ptr<Request> r = it->second;
        done_cv_.wait(lk, [&] { return r->done; });
        *ret = r->result;
        if (latency_ms) *latency_ms = r->latency_ms;
        // by key: another thread waiting for the same ticket may have
        // erased it already (it holds r, which stays valid)
        tickets_.erase(ticket);
        if (r->status) err = "request failed: " + fail_msg_;
        return r->status;
    }

    void stats(uint64_t out[6]) {
        std::lock_guard<std::mutex> lk(mu_);
        out[0] = submitted_;
        out[1] = completed_;
        out[2] = batches_;
        out[3] = max_inflight_;
        out[4] = peak_used_;
        out[5] = budget_;
    }

private:
    struct Base {
        void* ptr = nullptr;
        size_t n = 0;
    };

    // device bytes of a request: its inputs plus the share of a batch
    // workspace a job of its size adds
    size_t footprint(const Request& r) const {
        Layout L = make_layout(std::vector<JobSpec>(1, make_spec(nullptr, nullptr, r.n, r.nbits, 0, true)));
        size_t in = r.n * ((r.nbits + 7) / 8) + 64;
        if (r.base_id < 0) in += r.n * sizeof(blst_p1_affine);
        return in + workspace_bytes(L);
    }

    // under mu_: admit in arrival order while the budget allows; the copies
    // are issued by the copier thread (a copy from pageable host memory can
    // block the issuing thread for its whole duration)
    // A request is admitted when its footprint fits the HBM budget and the
    // points admitted but not finished stay within a few batches: requests
    // beyond that would only hold input buffers (and make the pool allocate)
    // long before a slot could take them.
    void admit() {
        const size_t lookahead = (kSlots + 1) * max_batch_points_;
        while (!pending_.empty()) {
            auto r = pending_.front();
            const bool idle = inflight_reqs_ == 0 && admitted_.empty();
            if (!idle && (used_ + r->need > budget_ || open_points_ + r->n > lookahead)) break;
            pending_.pop_front();
            used_ += r->need;
            open_points_ += r->n;
            peak_used_ = std::max(peak_used_, used_);
            HIPCK(hipEventCreateWithFlags(&r->copied, hipEventDisableTiming));
            admitted_.push_back(r);
            to_copy_.push_back(r);
            copy_cv_.notify_one();
        }
    }

    void copy_loop() {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            copy_cv_.wait(lk, [&] { return stop_ || !to_copy_.empty(); });
            if (to_copy_.empty()) {
                if (stop_) return;
                continue;
            }
            auto r = to_copy_.front();
            to_copy_.pop_front();
            r->copying = true;
            lk.unlock();
            bool ok = true;
            const size_t sbytes = r->n * ((r->nbits + 7) / 8);
            // device buffers come from the pool here, off the scheduler thread
            // (a new slab is a hipMalloc)
            try {
                r->d_scalars = pool_.get(sbytes + 16, r->d_scalars_cap);
                if (r->base_id < 0) r->d_points = pool_.get(r->n * sizeof(blst_p1_affine), r->d_points_cap);
            } catch (const HipFail&) {
                ok = false;
            }
            ok = ok && hipMemcpyAsync(r->d_scalars, r->host_scalars, sbytes, hipMemcpyHostToDevice,
                                      copy_stream_) == hipSuccess;
            if (r->base_id < 0)
                ok = ok && hipMemcpyAsync(r->d_points, r->host_points, r->n * sizeof(blst_p1_affine),
                                          hipMemcpyHostToDevice, copy_stream_) == hipSuccess;
            ok = ok && hipEventRecord(r->copied, copy_stream_) == hipSuccess;
            lk.lock();
            r->copying = false;
            r->copy_issued = true;
            if (r->failed) {
                // fail_all ran while this request was being copied and left
                // it here: its copies finish before its buffers go back to
                // the pool, and its host buffers are no longer read once
                // the waiter is released
                (void)hipStreamSynchronize(copy_stream_);
                (void)hipGetLastError();
                release(*r);
                r->done = true;
                done_cv_.notify_all();
            } else if (!ok) {
                copy_failed_ = true;
            }
            cv_.notify_all();
        }
    }

    // under mu_: whether a request's inputs are on the device
    bool landed(Request& r) {
        if (!r.landed && r.copy_issued && hipEventQuery(r.copied) == hipSuccess) r.landed = true;
        return r.landed;
    }

    // under mu_: points of the leading admitted requests that could form a
    // batch now (arrival order, up to one batch)
    size_t ready_points() {
        size_t pts = 0;
        for (auto& r : admitted_) {
            if (!landed(*r) || pts >= max_batch_points_) break;
            pts += r->n;
        }
        return pts;
    }

    // one batch of copied requests into a free slot: chosen under mu_, set
    // up and enqueued with mu_ released (the copier and submitters keep
    // going; only this thread touches the slot)
    void launch(Slot& s, std::unique_lock<std::mutex>& lk) {
        std::vector<std::shared_ptr<Request>> batch;
        std::vector<JobSpec> specs;
        size_t pts = 0;
        for (auto it = admitted_.begin(); it != admitted_.end();) {
            auto& r = *it;
            if (!landed(*r)) break;   // keep arrival order
            if (!batch.empty() && pts + r->n > max_batch_points_) break;
            specs.push_back(make_spec(r->base_id >= 0 ? bases_[r->base_id].ptr : r->d_points, r->d_scalars, r->n,
                                      r->nbits, 0, true));
            batch.push_back(r);
            pts += r->n;
            it = admitted_.erase(it);
        }
        if (batch.empty()) return;
        const size_t nreq = batch.size();
        s.reqs = std::move(batch);
        s.busy = true;
        inflight_reqs_ += nreq;
        batches_++;
        int inflight = 0;
        for (auto& o : slots_) inflight += o->busy ? 1 : 0;
        max_inflight_ = std::max<uint64_t>(max_inflight_, inflight);
        lk.unlock();
        try {
            s.L = make_layout(specs);
            ensure_workspace(s.ws, s.L);
            const DevLayout D = upload_layout(s.ws, s.L);
            const size_t ngroups = s.L.groups.size();
            if (s.host_res_cap < ngroups) {
                if (s.host_res) HIPCK(hipHostFree(s.host_res));
                s.host_res = nullptr;
                s.host_res_cap = 0;
                HIPCK(hipHostMalloc((void**)&s.host_res, std::max<size_t>(2 * ngroups, 64) * sizeof(hostec::xyzz6),
                                    hipHostMallocDefault));
                s.host_res_cap = std::max<size_t>(2 * ngroups, 64);
            }
            // window sums come back to the host; the folds run on the fold
            // pool when the batch completes, so the device moves on
            enqueue_batch(s.ws, s.L, D, true, s.host_res, false);
            HIPCK(hipEventRecord(s.done, s.ws.stream));
        } catch (...) {
            lk.lock();
            throw;
        }
        lk.lock();
    }

    // completion of a finished slot: window folds with mu_ released (only
    // this thread touches the slot), then the tickets under mu_
    void complete(Slot& s, std::unique_lock<std::mutex>& lk) {
        lk.unlock();
        fold_pool_.run(s.reqs.size(), [&s](size_t k) {
            const job_t& J = s.L.jobs[k];
            hostec::xyzz6 r;
            hostec::fold_windows(r, s.host_res + J.gbase, (int)J.nwin, (int)J.c);
            hostec::jac6 j;
            hostec::to_jac(j, r);
            std::memcpy(&s.reqs[k]->result, &j, sizeof(j));
        });
        lk.lock();
        const auto now = clock_t_::now();
        for (size_t k = 0; k < s.reqs.size(); k++) {
            auto& r = s.reqs[k];
            r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
            release(*r);
            r->done = true;
            completed_++;
        }
        inflight_reqs_ -= s.reqs.size();
        s.reqs.clear();
        s.busy = false;
        s.gpu_done = false;
        done_cv_.notify_all();
    }

    void release(Request& r) {
        if (r.d_scalars) pool_.put(r.d_scalars, r.d_scalars_cap);
        if (r.d_points) pool_.put(r.d_points, r.d_points_cap);
        r.d_scalars = r.d_points = nullptr;
        if (r.copied) (void)hipEventDestroy(r.copied);
        r.copied = nullptr;
        used_ -= r.need;
        open_points_ -= r.n;
    }

    // Under mu_: every request not yet completed fails.  Copies and batches
    // still queued on the device finish first (their buffers then go back to
    // the pool); a request the copier is working on is marked and released
    // by the copier once its copies are issued.
    void fail_all(const std::string& msg) {
        fail_msg_ = msg;
        (void)hipStreamSynchronize(copy_stream_);
        (void)hipStreamSynchronize(compute_stream_);
        (void)hipGetLastError();
        auto fail = [&](std::shared_ptr<Request>& r) {
            r->status = -1;
            if (r->copying) {
                r->failed = true;
                return;
            }
            release(*r);
            r->done = true;
        };
        for (auto& r : pending_) { r->status = -1; r->done = true; }
        pending_.clear();
        for (auto& r : admitted_) fail(r);
        admitted_.clear();
        to_copy_.clear();
        copy_failed_ = false;
        for (auto& s : slots_) {
            for (auto& r : s->reqs) fail(r);
            s->reqs.clear();
            s->busy = false;
            s->gpu_done = false;
        }
        inflight_reqs_ = 0;
        done_cv_.notify_all();
    }

    void loop() {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            try {
                if (copy_failed_) throw HipFail{"host to device copy failed"};
                bool progress = false;
                int on_device = 0;
                for (auto& s : slots_) {
                    if (s->busy && !s->gpu_done && hipEventQuery(s->done) == hipSuccess) s->gpu_done = true;
                    on_device += (s->busy && !s->gpu_done) ? 1 : 0;
                }
                admit();
                // Batch formation: an idle device takes whatever has landed;
                // behind a running batch, the next one is queued only once it
                // holds a quarter of a full batch, so requests arriving under
                // moderate load gather into batches large enough to amortise
                // the per-batch stages; never more than kQueued on the device.
                // Launching comes before completing, so the device has work
                // while finished batches are folded.
                for (auto& s : slots_) {
                    if (s->busy || admitted_.empty() || on_device >= kQueued) continue;
                    const size_t ready = ready_points();
                    if (ready == 0 || (on_device > 0 && ready < max_batch_points_ / 4)) break;
                    launch(*s, lk);
                    if (s->busy) {
                        on_device++;
                        progress = true;
                    }
                }
                for (auto& s : slots_) {
                    if (s->busy && s->gpu_done) {
                        complete(*s, lk);
                        progress = true;
                    }
                }
                const bool busy = inflight_reqs_ > 0 || !admitted_.empty();
                if (stop_ && !busy && pending_.empty()) break;
                if (!progress) {
                    if (busy) cv_.wait_for(lk, std::chrono::microseconds(50));
                    else cv_.wait(lk, [&] { return stop_ || !pending_.empty(); });
                }
            } catch (const HipFail& f) {
                fail_all(f.msg);
            } catch (const std::exception& e) {
                fail_all(e.what());
            }
        }
    }

    int device_ = 0;
    size_t budget_ = 0, used_ = 0, peak_used_ = 0, max_batch_points_;
    size_t open_points_ = 0;   // points of admitted requests not yet finished
    hipStream_t copy_stream_ = nullptr;
    hipStream_t compute_stream_ = nullptr;   // shared by the slots
    FoldPool fold_pool_;
    std::unique_ptr<Slot> slots_[kSlots];
    std::vector<Base> bases_;
    BufferPool pool_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::shared_ptr<Request>> pending_;
    std::deque<std::shared_ptr<Request>> admitted_;
    std::deque<std::shared_ptr<Request>> to_copy_;
    std::condition_variable copy_cv_;
    bool copy_failed_ = false;
    std::thread copier_;
    std::map<int64_t, std::shared_ptr<Request>> tickets_;
    int64_t next_ticket_ = 0;
    size_t inflight_reqs_ = 0;
    uint64_t submitted_ = 0, completed_ = 0, batches_ = 0, max_inflight_ = 0;
    std::string fail_msg_;
    bool stop_ = false;
    std::thread worker_;
};

}  // namespace host
}  // namespace bls

struct blst_hip_server {
    std::unique_ptr<bls::host::Server> impl;
};

extern "C" {

blst_hip_server* blst_hip_server_create(size_t hbm_budget_bytes, size_t max_batch_points) {
    try {
        auto s = std::make_unique<blst_hip_server>();
        s->impl = std::make_unique<bls::host::Server>(hbm_budget_bytes,
                                                      max_batch_points ? max_batch_points : ((size_t)1 << 24));
        bls::host::g_err.clear();
        return s.release();
    } catch (const bls::host::HipFail& f) {
        bls::host::g_err = f.msg;
    } catch (const std::exception& e) {
        bls::host::g_err = e.what();
    } catch (...) {
        bls::host::g_err = "server creation failed";
    }
    return nullptr;
}

void blst_hip_server_destroy(blst_hip_server* s) {
    try {
        delete s;
    } catch (...) {
        bls::host::g_err = "server shutdown failed";
    }
}

int blst_hip_server_add_bases(blst_hip_server* s, const blst_p1_affine* points, size_t npoints) {
    if (!s || !points || npoints == 0) { bls::host::g_err = "invalid arguments"; return -1; }
    try {
        const int id = s->impl->add_bases(points, npoints);
        bls::host::g_err.clear();
        return id;
    } catch (const bls::host::HipFail& f) {
        bls::host::g_err = f.msg;
    } catch (const std::exception& e) {
        bls::host::g_err = e.what();
    } catch (...) {
        bls::host::g_err = "add_bases failed";
    }
    return -1;
}

int64_t blst_hip_server_submit(blst_hip_server* s, int base_id, const blst_p1_affine* points,
                               const blst_byte* scalars, size_t npoints, size_t nbits) {
    if (!s || !scalars || npoints == 0 || npoints >= 0x7fffffffULL || nbits == 0 || nbits > 256 ||
        (base_id < 0 && (!points || ((uintptr_t)points % 8) != 0))) {
        bls::host::g_err = "invalid arguments";
        return -1;
    }
    try {
        std::string err;
        const int64_t t = s->impl->submit(base_id, points, scalars, npoints, nbits, err);
        if (t < 0) bls::host::g_err = "request rejected: " + err;
        else bls::host::g_err.clear();
        return t;
    } catch (const bls::host::HipFail& f) {
        bls::host::g_err = f.msg;
    } catch (const std::exception& e) {
        bls::host::g_err = e.what();
    } catch (...) {
        bls::host::g_err = "submit failed";
    }
    return -1;
}

int blst_hip_server_wait(blst_hip_server* s, int64_t ticket, blst_p1* ret, double* latency_ms) {
    if (!s || !ret) { bls::host::g_err = "invalid arguments"; return -1; }
    try {
        std::string err;
        const int rc = s->impl->wait(ticket, ret, latency_ms, err);
        if (rc) bls::host::g_err = err;
        else bls::host::g_err.clear();
        return rc;
    } catch (...) {
        bls::host::g_err = "wait failed";
    }
    return -1;
}

void blst_hip_server_stats(blst_hip_server* s, uint64_t out[6]) {
    try {
        if (s) s->impl->stats(out);
    } catch (...) {
        bls::host::g_err = "stats failed";
    }
}

}  // extern "C"
