// Host-side pipeline of the gfx950 MSM backend, shared by the C ABI entry
// points, the captured batch plans and the serving loop: device buffers,
// batch layouts and the enqueue of the device stages.
#pragma once
#include "launch.h"
#include "host_ec.h"

#include <atomic>
#include <string>
#include <vector>

namespace bls {
namespace host {

extern thread_local std::string g_err;
extern thread_local double g_phase[7];

struct HipFail {
    std::string msg;
};

#define HIPCK(expr)                                                                   \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw HipFail{std::string(#expr) + ": " + hipGetErrorString(e_)};         \
    } while (0)

// Device allocations larger than this many bytes fail as a device out of
// memory would (0: no cap).  Only the test kit sets it, to drive the
// fail-closed paths of the entry points from the host.
extern std::atomic<size_t> g_alloc_cap;
// The staging thread of a split host-buffer MSM fails before this point
// range (-1: never); only the test kit sets it.
extern std::atomic<int> g_stage_fail_range;

inline void device_malloc(void** p, size_t bytes) {
    const size_t cap = g_alloc_cap.load(std::memory_order_relaxed);
    if (cap && bytes > cap)
        throw HipFail{"device allocation of " + std::to_string(bytes) + " bytes refused (allocation cap)"};
    HIPCK(hipMalloc(p, bytes));
}

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool fixed = false;           // a region of memory this buffer does not own
    std::vector<void*> retired;   // outgrown allocations, freed with the buffer
    template <class T> T* as() const { return static_cast<T*>(p); }
    // use [q, q + bytes) of memory owned elsewhere; ensure() then never allocates
    void attach(void* q, size_t bytes) {
        p = q;
        cap = bytes;
        fixed = true;
    }
    void ensure(size_t bytes) {
        if (bytes <= cap && p) return;
        if (fixed) throw HipFail{"fixed device buffer too small"};
        // hipFree waits for the whole device, which would drain work queued
        // behind this call: an outgrown buffer is kept until destruction
        // instead (geometric growth bounds what is kept to the current size)
        const size_t old_cap = p ? cap : 0;
        if (p) retired.push_back(p);
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(std::max<size_t>(bytes + bytes / 8, 2 * old_cap), 4096);
        device_malloc(&p, want);
        cap = want;
    }
    ~DevBuf() {
        for (void* q : retired) (void)hipFree(q);
        if (p && !fixed) (void)hipFree(p);
    }
};

// Device memory and stream of one in-flight batch.  The per-thread context,
// each captured batch plan and each serving slot own one.
struct Workspace {
    hipStream_t stream = nullptr;
    DevBuf ent, rank, counts, hist, off, sorted, buckets, slot_key, slots, heavy, H, scan_tmp, W, res, tables;
    DevBuf cnt_sum;                       // bucket entry counts summed over point ranges
    uint8_t* tables_pinned = nullptr;     // page-locked staging of the layout tables
    size_t tables_cap = 0;
    hipEvent_t ev[8] = {};
    void* pinned = nullptr;               // page-locked result staging
    size_t pinned_cap = 0;
    bool owns_stream = false;
    // layout tables in and window sums out by a copy kernel through the
    // device mapping of the page-locked buffers instead of copy-engine
    // transfers (the serving loop's batches: an engine transfer queued behind
    // a running batch holds up the request copies of the copy stream)
    bool kernel_transfers = false;
    // own stream, or a stream shared with other workspaces (kernels of
    // different workspaces then run one after another)
    void init(hipStream_t shared = nullptr) {
        if (shared) {
            stream = shared;
        } else {
            HIPCK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            owns_stream = true;
        }
        for (auto& e : ev) HIPCK(hipEventCreate(&e));
    }
    // staging for the layout tables: page-locked, so their upload is a true
    // asynchronous copy that does not wait for work queued on the stream
    // fixed host staging (set by the owner, never reallocated): host_fixed
    // serves both the layout tables and the results, one after the other
    uint8_t* host_fixed = nullptr;
    size_t host_fixed_cap = 0;
    uint8_t* tables_stage(size_t bytes) {
        if (host_fixed) {
            if (bytes > host_fixed_cap) throw HipFail{"fixed host staging too small"};
            return host_fixed;
        }
        if (bytes > tables_cap) {
            if (tables_pinned) HIPCK(hipHostFree(tables_pinned));
            tables_pinned = nullptr;
            tables_cap = 0;
            const size_t want = std::max<size_t>(2 * bytes, 1 << 16);
            HIPCK(hipHostMalloc((void**)&tables_pinned, want, hipHostMallocDefault));
            tables_cap = want;
        }
        return tables_pinned;
    }
    void* host_buf(size_t bytes) {
        if (host_fixed) {
            if (bytes > host_fixed_cap) throw HipFail{"fixed host staging too small"};
            return host_fixed;
        }
        if (bytes > pinned_cap) {
            if (pinned) HIPCK(hipHostFree(pinned));
            pinned = nullptr;
            pinned_cap = 0;
            HIPCK(hipHostMalloc(&pinned, bytes, hipHostMallocDefault));
            pinned_cap = bytes;
        }
        return pinned;
    }
    ~Workspace() {
        if (pinned) (void)hipHostFree(pinned);
        if (tables_pinned) (void)hipHostFree(tables_pinned);
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
        if (stream && owns_stream) (void)hipStreamDestroy(stream);
    }
};

// One MSM of a batch as the host sees it.
struct JobSpec {
    const void* points;    // device, n affine points
    const void* scalars;   // device, n * ceil(nbits/8) bytes
    uint32_t n, nbits, c;
    uint32_t T, L;         // tile shape (the batch uses the T of its largest job)
    uint32_t bit_base;     // tile mode: first bit of the single window
    bool tile;             // one window at bit_base, no window fold
    uint32_t win0 = 0;     // window range [win0, win0 + nwin) of the job
    uint32_t nwin = 0;     // (0: all windows)
    bool g2 = false;       // G2 points (Fp2 coordinates, 192-byte affine)
};

// Tile shape and window of one MSM size: c window bits, T sorted entries
// per accumulate lane, L buckets per reduction lane.
struct Tune {
    uint32_t c, T, L;
};

// Tuned parameters for an MSM of n points (batch: as one job of a batch).
// Entries come from the autotuner's table when loaded, else from a cost model.
Tune tuned(size_t n, size_t nbits, bool batch);

// Key, entry, point, group and segment ranges of every job of a batch.
struct Layout {
    std::vector<job_t> jobs;
    std::vector<uint32_t> pbase, ebase, kbase;
    std::vector<group_t> groups;
    std::vector<uint32_t> sbase;
    std::vector<sgroup_t> sgroups;     // sort tiles per group
    std::vector<uint32_t> gkb;         // first key of every group
    // sort tiles (group, chunk), ordered by LDS size class; class k covers
    // tiles [tile_begin[k], tile_begin[k + 1]) and needs tile_lds[k] bytes
    std::vector<uint2> tiles;
    std::vector<uint32_t> tile_begin;
    std::vector<size_t> tile_lds;
    uint32_t total_points = 0, total_entries = 0, nkeys = 0, total_segments = 0, max_n = 0;
    uint32_t ntiles = 0, max_B = 0;
    size_t count_words = 0;
    // binned sort (every window of 2^6 .. 2^15 buckets, at least half of the
    // entries in windows of 2^12 or more): sort groups hold tiles of
    // kBinTile points and bin-major counts, one tile class
    bool binned = false;
    // reduction tree: per pass, blocks and node offsets (in nodes of two
    // xyzz) of its input and output arrays; per pass and group the first
    // block and (first input node, input nodes, first output node)
    std::vector<uint32_t> tree_blocks;
    std::vector<size_t> tree_in, tree_out;
    std::vector<uint32_t> tree_blk;
    std::vector<uint4> tree_tn;
    size_t tree_nodes = 0;
    uint32_t T = 32;       // sorted entries per accumulate lane
    bool g2 = false;       // every job of the batch is a G2 MSM
    size_t point_bytes() const { return g2 ? 384 : 192; }   // one XYZZ point
};

// Device copies of the layout tables (one upload per batch).
struct DevLayout {
    const job_t* jobs;
    const uint32_t *pbase, *ebase, *kbase, *sbase, *gkb;
    const uint2* tiles;
    const group_t* groups;
    const sgroup_t* sgroups;
    const uint32_t* tree_blk;
    const uint4* tree_tn;
};

uint32_t device_window(size_t n, size_t nbits);
JobSpec make_spec(const void* points, const void* scalars, size_t n, size_t nbits, size_t window, bool batch);
Layout make_layout(const std::vector<JobSpec>& specs);
// bytes of a layout's device tables; upload_layout writes them at byte
// offset base of the workspace's table buffer (several layouts in flight on
// one workspace take separate regions, sized beforehand)
size_t layout_table_bytes(const Layout& L);
DevLayout upload_layout(Workspace& cx, const Layout& L, size_t base = 0);
void ensure_workspace(Workspace& cx, const Layout& L);
// Bytes of device memory that carve_workspace takes for a layout (its
// buffers and layout tables, each 256-byte aligned).
size_t carved_bytes(const Layout& L);
// Points the workspace's buffers (and its layout-table buffer) into
// [base, base + cap) with the sizes ensure_workspace would allocate; throws
// if the region is too small.  Nothing is allocated.
void carve_workspace(Workspace& cx, const Layout& L, uint8_t* base, size_t cap);
size_t workspace_bytes(const Layout& L);
bool tiled_sort(const Layout& L);
// One point range of an MSM whose points arrive in ranges: prev holds the
// bucket entry counts of the earlier ranges (nullptr for the first), cnt_sum
// receives the running counts, and only the last range reduces.
struct RangeRole {
    const uint32_t* prev = nullptr;
    uint32_t* cnt_sum = nullptr;
    bool reduce = true;
};
void enqueue_batch(Workspace& cx, const Layout& L, const DevLayout& D, bool host_fold, void* host_out,
                   bool timing = true, const RangeRole& range = RangeRole());
void record_phases(Workspace& cx);

}  // namespace host
}  // namespace bls
